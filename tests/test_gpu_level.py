"""GPU parity of the level-step solve (CAP_ALGO_LEVEL, k_level) against the oracle.

One workgroup walks the levels; every row's sum is the sequential stored-order
chain, so x must equal the oracle's fused solve bit for bit (and, with
arith="reference", the reference executor's unfused one), in FP64 and FP32.
The plan accepts a matrix only when its ring and padding rules hold; those
rules are restated here from the levels and checked against the plan's
answer, so a refusal of an acceptable matrix fails as well as a wrong x.
"""
import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import api, _native
from conftest import random_lower, rel_maxnorm, upper_from_lower, drop_diagonal

pytestmark = pytest.mark.gpu

RING_BYTES = 65536  # the plan's LDS ring limit


def accepts(rp, ci, vbytes):
    """The plan's acceptance rule, from the oracle's row levels: every
    off-diagonal entry c of a row in level L satisfies c > M_L - R (M_L the
    largest row index in levels <= L, R a power of two >= 64 within the
    ring's bytes), and the levels padded to whole steps of B rows hold at
    most max(4 n, 65536) slots, B the smallest of 64, 128, 192, 256 that
    needs at most 1/8 more steps than 256."""
    n = len(rp) - 1
    lv = np.asarray(oracle.row_levels(n, rp, ci))
    L = int(lv.max())
    width = np.bincount(lv, minlength=L + 1)
    top = np.full(L + 1, -1)
    np.maximum.at(top, lv, np.arange(n))
    top = np.maximum.accumulate(top)
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = ci != rows
    need = int((top[lv[rows[off]]] - ci[off] + 1).max()) if off.any() else 1
    ring = 64
    while ring < need:
        ring *= 2
    def steps(B):
        return int(np.sum((width[1:] + B - 1) // B))
    B = min(B for B in (64, 128, 192, 256) if steps(B) <= steps(256) + steps(256) // 8)
    return ring * vbytes <= RING_BYTES and steps(B) * B <= max(4 * n, 65536)


def _plan(cuda, rp, ci, v, **kw):
    import torch
    from capellinisptrsv_amd import Plan
    return Plan(torch.from_numpy(rp).to(cuda), torch.from_numpy(ci).to(cuda), torch.from_numpy(np.ascontiguousarray(v)).to(cuda), **kw)


def _run(plan, cuda, b):
    import torch
    x = plan.solve(torch.from_numpy(np.ascontiguousarray(b)).to(cuda))
    torch.cuda.synchronize()
    return x.cpu().numpy().reshape(-1)


def _mats():
    yield "cfg0", api.make_config(0, seed=0)                        # 723 levels of ~138 rows
    yield "band5k", api.generate("band", n=5000, offdiag=3, band=300, seed=3)
    yield "tiny", api.generate("band", n=50, offdiag=3, band=10, seed=3)
    yield "one_row", api.generate("band", n=1, offdiag=3, band=10, seed=3)
    # rows of 7 and 12 off-diagonals (8 carried per row, the rest from the CSR),
    # levels wider than one step
    yield "level_o6", api.generate("level", n=6000, levels=20, width=300, depth=2, offdiag=6, seed=2)
    yield "level_o12", api.generate("level", n=20000, levels=40, width=500, depth=2, offdiag=12, seed=5)
    yield "dup", random_lower(4000, max_off=6, band=300, seed=2, dup=True)
    # a few rows per level: one-wavefront workgroups
    yield "narrow_band", random_lower(3000, max_off=5, band=40, seed=1)


MATS = list(_mats())


@pytest.mark.parametrize("name,mat", MATS, ids=[m[0] for m in MATS])
def test_level_bit_exact(cuda, name, mat):
    rp, ci, v = mat
    n = len(rp) - 1
    assert accepts(rp, ci, 8), name  # every case here is one the plan must take
    b = api.uniform(11, n)
    plan = _plan(cuda, rp, ci, v, algo="level")
    info = plan.info()
    assert info["algo"] == "level" and info["row_threads"] in (64, 128, 192, 256)
    x = _run(plan, cuda, b)
    assert plan.status() == 0
    assert np.array_equal(x.view(np.uint64), oracle.csr_forward(rp, ci, v, b, fused=True).view(np.uint64)), name
    # again on the same plan (no per-solve state to reset)
    b2 = api.uniform(12, n)
    assert np.array_equal(_run(plan, cuda, b2), oracle.csr_forward(rp, ci, v, b2, fused=True)), name
    plan.close()
    # the reference executor's arithmetic: bit-identical to it
    plan = _plan(cuda, rp, ci, v, algo="level", arith="reference")
    x = _run(plan, cuda, b)
    assert np.array_equal(x.view(np.uint64), oracle.csr_forward(rp, ci, v, b, fused=False).view(np.uint64)), name
    plan.close()


@pytest.mark.parametrize("name,mat", MATS[:3] + MATS[4:6], ids=[m[0] for m in MATS[:3] + MATS[4:6]])
def test_level_fp32(cuda, name, mat):
    rp, ci, v = mat
    n = len(rp) - 1
    v32 = v.astype(np.float32)
    b = api.uniform(13, n)
    b32 = b.astype(np.float32)
    plan = _plan(cuda, rp, ci, v32, algo="level")
    x = _run(plan, cuda, b32)
    assert np.array_equal(x.view(np.uint32), oracle.csr_forward(rp, ci, v32, b32, fused=True).view(np.uint32)), name
    assert rel_maxnorm(x.astype(np.float64), oracle.csr_forward(rp, ci, v, b, fused=True)) <= 1e-5
    plan.close()


def test_level_refusals(cuda):
    """Matrices outside the rules are refused with CAP_ERR_UNSUPPORTED at
    plan creation: a ring beyond 64 KiB (wide levels reaching far back), too
    many padded slots (many narrow levels); an explicit level plan refuses
    several right-hand sides."""
    import torch
    cases = [("wide_levels", api.make_config(1, seed=0, scale=0.05)),
             ("deep_band", api.make_config(2, seed=0, scale=0.02)),
             ("deep_band_2", api.make_config(2, seed=1, scale=0.04))]
    for name, (rp, ci, v) in cases:
        assert not accepts(rp, ci, 8), name
        with pytest.raises(RuntimeError, match=str(_native.CAP_ERR_UNSUPPORTED)):
            _plan(cuda, rp, ci, v, algo="level")
    # FP32 halves the ring's bytes: a matrix refused in FP64 for its ring
    rp, ci, v = api.generate("level", n=20000, levels=20, width=1000, depth=3, offdiag=12, seed=1)
    assert not accepts(rp, ci, 8) and accepts(rp, ci, 4)
    with pytest.raises(RuntimeError, match=str(_native.CAP_ERR_UNSUPPORTED)):
        _plan(cuda, rp, ci, v, algo="level")
    v32 = v.astype(np.float32)
    plan = _plan(cuda, rp, ci, v32, algo="level")
    b = api.uniform(14, len(rp) - 1).astype(np.float32)
    assert np.array_equal(_run(plan, cuda, b), oracle.csr_forward(rp, ci, v32, b, fused=True))
    with pytest.raises(RuntimeError, match=str(_native.CAP_ERR_UNSUPPORTED)):
        plan.solve(torch.zeros(len(rp) - 1, 2, dtype=torch.float32, device=cuda))
    plan.close()


@pytest.mark.parametrize("upper,unit", [(True, False), (False, True), (True, True)])
def test_level_triangular_kinds(cuda, upper, unit):
    rp, ci, v = api.generate("band", n=5000, offdiag=3, band=300, seed=3)
    n = len(rp) - 1
    b = api.uniform(15, n)
    if upper:
        rp2, ci2, v2 = upper_from_lower(rp, ci, v)
    else:
        rp2, ci2, v2 = rp, ci, v
    if unit:
        rp2, ci2, v2 = drop_diagonal(rp2, ci2, v2, upper=upper)
    plan = _plan(cuda, rp2, ci2, v2, algo="level", upper=upper, unit_diagonal=unit)
    x = _run(plan, cuda, b)
    ref = oracle.csr_solve(rp2, ci2, v2, b, upper=upper, unit=unit, fused=True)
    assert np.array_equal(x, ref)
    plan.close()


def test_level_autotuned_plan(cuda):
    """On the 100k band (723 levels of ~138 rows) the autotuner times the
    level steps with the other candidates; whatever it picks is bit-exact,
    its several-RHS solves run the write-first kernel bit-exact, and its
    refined solve converges to 1e-12."""
    import torch
    rp, ci, v = api.make_config(0, seed=0)
    n = len(rp) - 1
    plan = _plan(cuda, rp, ci, v, algo="auto")
    b = api.uniform(16, n)
    assert np.array_equal(_run(plan, cuda, b), oracle.csr_forward(rp, ci, v, b, fused=True))
    k = 3
    bk = api.uniform(17, n * k).reshape(n, k)
    xk = plan.solve(torch.from_numpy(bk).to(cuda))
    torch.cuda.synchronize()
    xk = xk.cpu().numpy()
    for q in range(k):
        assert np.array_equal(xk[:, q], oracle.csr_forward(rp, ci, v, bk[:, q].copy(), fused=True)), q
    xr, _, _ = plan.solve_refined(torch.from_numpy(b).to(cuda), tol=1e-12)
    torch.cuda.synchronize()
    assert rel_maxnorm(xr.cpu().numpy(), oracle.csr_forward(rp, ci, v, b, fused=True)) <= 1e-12
    plan.close()


def test_level_refined_and_adjoint(cuda):
    """An explicit level plan refines (its FP32 sweeps run the thread-per-row
    kernels) and differentiates like any other plan."""
    import torch
    rp, ci, v = api.generate("band", n=5000, offdiag=3, band=300, seed=3)
    n = len(rp) - 1
    plan = _plan(cuda, rp, ci, v, algo="level")
    b = api.uniform(18, n)
    ref = oracle.csr_forward(rp, ci, v, b, fused=True)
    xr, _, _ = plan.solve_refined(torch.from_numpy(b).to(cuda), tol=1e-12)
    torch.cuda.synchronize()
    assert rel_maxnorm(xr.cpu().numpy(), ref) <= 1e-12
    x = torch.from_numpy(ref).to(cuda)
    g = api.uniform(19, n)
    gb, _ = plan.adjoint(torch.from_numpy(g).to(cuda), x)
    torch.cuda.synchronize()
    assert rel_maxnorm(gb.cpu().numpy(), oracle.adjoint(rp, ci, v, g, fused=True)) <= 1e-12
    plan.close()


def test_level_graph_replay(cuda):
    """A level solve captured into a graph replays with fresh b each time."""
    import torch
    rp, ci, v = api.generate("band", n=5000, offdiag=3, band=300, seed=3)
    n = len(rp) - 1
    plan = _plan(cuda, rp, ci, v, algo="level")
    bt = torch.zeros(n, dtype=torch.float64, device=cuda)
    xt = torch.zeros(n, dtype=torch.float64, device=cuda)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.solve(bt, out=xt, stream=s)  # warm-up outside the capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.solve(bt, out=xt, stream=torch.cuda.current_stream())
    for seed in (21, 22, 23):
        b = api.uniform(seed, n)
        bt.copy_(torch.from_numpy(b))
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(xt.cpu().numpy(), oracle.csr_forward(rp, ci, v, b, fused=True)), seed
    plan.close()


@pytest.mark.parametrize("k", [2, 4])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_level_several_rhs(cuda, k, dtype):
    """Rows of at most 4 off-diagonals take 2 or 4 right-hand sides per solve
    (one thread per row, the ring K values wide), each column bit-exact;
    3 right-hand sides, or rows past 4 entries, are refused."""
    import torch
    for name, (rp, ci, v) in MATS[:4]:
        n = len(rp) - 1
        vv = v.astype(dtype)
        bk = api.uniform(30 + k, n * k).astype(dtype).reshape(n, k)
        plan = _plan(cuda, rp, ci, vv, algo="level")
        x = plan.solve(torch.from_numpy(np.ascontiguousarray(bk)).to(cuda))
        torch.cuda.synchronize()
        x = x.cpu().numpy().reshape(n, k)
        for q in range(k):
            assert np.array_equal(x[:, q], oracle.csr_forward(rp, ci, vv, bk[:, q].copy(), fused=True)), (name, q)
        with pytest.raises(RuntimeError, match=str(_native.CAP_ERR_UNSUPPORTED)):
            plan.solve(torch.zeros(n, 3, dtype=torch.from_numpy(vv).dtype, device=cuda))
        plan.close()
    rp, ci, v = MATS[6][1]  # rows of up to 7 off-diagonals: 8 carried, one right-hand side only
    plan = _plan(cuda, rp, ci, v.astype(dtype), algo="level")
    with pytest.raises(RuntimeError, match=str(_native.CAP_ERR_UNSUPPORTED)):
        plan.solve(torch.zeros(len(rp) - 1, k, dtype=torch.from_numpy(v.astype(dtype)).dtype, device=cuda))
    plan.close()


def test_level_several_rhs_reference_arith_and_autotuned(cuda):
    """Two right-hand sides in the reference arithmetic match the reference
    executor's bits; an autotuned plan on the 100k band solves 2 and 4
    right-hand sides bit-exact (level steps or write-first, whichever it runs)."""
    import torch
    rp, ci, v = MATS[1][1]
    n = len(rp) - 1
    b2 = api.uniform(40, n * 2).reshape(n, 2)
    plan = _plan(cuda, rp, ci, v, algo="level", arith="reference")
    x = plan.solve(torch.from_numpy(b2).to(cuda)).cpu().numpy()
    for q in range(2):
        assert np.array_equal(x[:, q], oracle.csr_forward(rp, ci, v, b2[:, q].copy(), fused=False)), q
    plan.close()
    rp, ci, v = MATS[0][1]
    n = len(rp) - 1
    plan = _plan(cuda, rp, ci, v, algo="auto")
    for k in (2, 4):
        bk = api.uniform(41 + k, n * k).reshape(n, k)
        x = plan.solve(torch.from_numpy(bk).to(cuda)).cpu().numpy()
        for q in range(k):
            assert np.array_equal(x[:, q], oracle.csr_forward(rp, ci, v, bk[:, q].copy(), fused=True)), (k, q)
    plan.close()
