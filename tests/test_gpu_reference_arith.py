"""GPU solves in the reference arithmetic (cap_opts.arith = CAP_ARITH_REFERENCE).

Each term of a row's sum is added as the reference's serial executor adds it
(openCL/SynFree_csc/sptrsv_syncfree_serialref.h:67-68: left_sum += x*val,
the product rounded, then the sum; g++ -O3 without FMA contraction), in the
same ascending-column order.  Bar: bit-identical to that executor -- the
reference's own code compiled from its sources into oracle/_ref where that
build is present, and to the oracle's unfused forward substitution (pinned
bit-equal to it on the CPU) everywhere -- for every algorithm that keeps the
sequential order: thread-per-row (both readiness protocols, LDS window,
128/256 threads), two-part, mix (thread and wavefront rows), one-workgroup,
batches and the gather adjoint (the executor's BACKWARD mode).  Cases: the
.mtx fixtures through change2tran, duplicates, the 100k band and the 2M
wide-level workload at full size, FP32 and FP64, 1-16 right-hand sides.
CSC and the partitioned solve refuse the mode.
"""
import os

import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import api
from conftest import random_lower

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "mtx")


def _executor(rp, ci, v, b, k=1):
    """The reference's serial executor (forward substitution) on L x = b."""
    n, nnz = len(rp) - 1, len(ci)
    cr, cp, cv = oracle.transpose(n, n, nnz, rp, ci, v)
    indeg = oracle.csc_analyser(cr, n, nnz)
    variant = "d64" if v.dtype == np.float64 else "s64"
    if oracle.ref_available(variant):
        return oracle.ref_csc_executor(cp, cr, cv, indeg, n, n, 0, k, b, variant=variant)
    return oracle.csc_executor(cp, cr, cv, indeg, n, n, 0, k, b)


def _plan(cuda, rp, ci, v, **kw):
    import torch
    from capellinisptrsv_amd import Plan
    return Plan(*(torch.from_numpy(a).to(cuda) for a in (rp, ci, v)), arith="reference", **kw)


def _solve(cuda, rp, ci, v, b, k=1, **kw):
    import torch
    plan = _plan(cuda, rp, ci, v, **kw)
    bt = torch.from_numpy(np.ascontiguousarray(b.reshape(-1, k) if k > 1 else b)).to(cuda)
    x = plan.solve(bt).cpu().numpy().reshape(-1)
    assert plan.status() == 0
    info = plan.info()
    plan.close()
    return x, info


def _fixture_lowers():
    for name in ("general_real.mtx", "symmetric_real.mtx", "integer_symmetric.mtx", "pattern_general.mtx",
                 "skew.mtx", "MixedCase_banner.mtx"):
        ret, m, n, nnz, rp, ci, v = oracle.read_mtx(os.path.join(GOLD, name))
        assert ret == 0
        mm = min(m, n)
        # real values kept (the reference's unit values make every sum exact)
        nnzL, rpL, ciL, vL = api.change2tran(mm, int(rp[mm]), rp[:mm + 1], ci, v, keep_values=True)
        yield name, (rpL, ciL, vL)


def _matrices():
    yield from _fixture_lowers()
    yield "band", random_lower(3000, max_off=5, band=40, seed=21)
    yield "dup", random_lower(2000, max_off=6, band=30, seed=22, dup=True)
    yield "cfg0", api.make_config(0, seed=0)
    yield "cfg2_small", api.make_config(2, seed=0, scale=0.02)


MATS = list(_matrices())
KERNELS = [dict(algo="wrtfst", sync="value"), dict(algo="wrtfst", sync="value", lds_window=True),
           dict(algo="wrtfst", sync="flag"), dict(algo="wrtfst", sync="value", row_threads=128),
           dict(algo="twopart", sync="value"), dict(algo="mix", sync="value", border=4),
           dict(algo="mix", sync="value", border=1_000_000), dict(algo="local"), dict(algo="auto")]


@pytest.mark.parametrize("kw", KERNELS, ids=lambda d: "-".join(f"{k}={v}" for k, v in d.items()))
@pytest.mark.parametrize("name,mat", MATS, ids=[m[0] for m in MATS])
def test_reference_arithmetic_bit_identical_to_serial_executor(cuda, kw, name, mat):
    rp, ci, v = mat
    n = len(rp) - 1
    b = api.uniform(41, n)
    x, info = _solve(cuda, rp, ci, v, b, **kw)
    want = _executor(rp, ci, v, b)
    assert np.array_equal(x.view(np.uint64), want.view(np.uint64)), (name, info)
    assert np.array_equal(want, oracle.csr_forward(rp, ci, v, b, fused=False))


def test_fused_and_reference_arithmetic_differ_only_in_rounding(cuda):
    """On the wide-level workload the fused default is not bit-identical to the
    executor but stays within a few ulps; the reference mode is exact."""
    rp, ci, v = api.make_config(1, seed=0, scale=0.05)
    b = api.uniform(42, len(rp) - 1)
    want = _executor(rp, ci, v, b)
    x_ref, _ = _solve(cuda, rp, ci, v, b, algo="wrtfst")
    import torch
    from capellinisptrsv_amd import Plan
    plan = Plan(*(torch.from_numpy(a).to(cuda) for a in (rp, ci, v)), algo="wrtfst")
    x_fused = plan.solve(torch.from_numpy(b).to(cuda)).cpu().numpy()
    plan.close()
    assert np.array_equal(x_ref, want)
    assert not np.array_equal(x_fused, want)
    assert np.max(np.abs(x_fused - want)) <= 1e-14 * np.max(np.abs(want))


def test_wide_level_2m_full_size(cuda):
    rp, ci, v = api.make_config(1, seed=0)
    b = api.uniform(1, len(rp) - 1)
    x, info = _solve(cuda, rp, ci, v, b, algo="auto")
    assert np.array_equal(x.view(np.uint64), _executor(rp, ci, v, b).view(np.uint64)), info


@pytest.mark.parametrize("name,mat", MATS[-4:], ids=[m[0] for m in MATS[-4:]])
def test_reference_arithmetic_fp32(cuda, name, mat):
    rp, ci, v = mat
    v32 = v.astype(np.float32)
    b = api.uniform(43, len(rp) - 1).astype(np.float32)
    for kw in (dict(algo="wrtfst", sync="value"), dict(algo="mix", sync="value", border=4), dict(algo="local")):
        x, info = _solve(cuda, rp, ci, v32, b, **kw)
        assert np.array_equal(x.view(np.uint32), _executor(rp, ci, v32, b).view(np.uint32)), (name, info)


@pytest.mark.parametrize("k", [2, 16])
@pytest.mark.parametrize("algo", ["wrtfst", "local"])
def test_reference_arithmetic_multi_rhs(cuda, k, algo):
    rp, ci, v = api.make_config(0, seed=3, scale=0.3)
    n = len(rp) - 1
    b = api.uniform(44, n * k)
    x, info = _solve(cuda, rp, ci, v, b, k=k, algo=algo)
    assert np.array_equal(x, _executor(rp, ci, v, b, k=k)), info


@pytest.mark.parametrize("algo", ["wrtfst", "local"])
def test_reference_arithmetic_batch(cuda, algo):
    import torch
    from capellinisptrsv_amd import Batch
    mats, row_off, rp, ci, v = api.make_batch(12, seed=4, n_hi=20_000)
    k = 4
    N = int(row_off[-1])
    b = api.uniform(45, N * k)
    bt = Batch(*(torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v)), k=k, algo=algo, arith="reference")
    x = bt.solve(torch.from_numpy(b.reshape(N, k)).to(cuda)).cpu().numpy().reshape(-1)
    bt.close()
    for m, (mrp, mci, mv) in enumerate(mats):
        lo, hi = row_off[m] * k, row_off[m + 1] * k
        assert np.array_equal(x[lo:hi], _executor(mrp, mci, mv, b[lo:hi], k=k)), m


def test_reference_arithmetic_gather_adjoint(cuda):
    """grad_b = L^{-T} g is the executor's BACKWARD mode on the same CSC."""
    import torch
    rp, ci, v = api.make_config(4, seed=5, scale=0.05)
    n, nnz = len(rp) - 1, len(ci)
    plan = _plan(cuda, rp, ci, v, algo="wrtfst", adjoint="gather")
    b, g = api.uniform(46, n), api.uniform(47, n)
    x = plan.solve(torch.from_numpy(b).to(cuda))
    gb, _ = plan.adjoint(torch.from_numpy(g).to(cuda), x)
    gb = gb.cpu().numpy()
    assert plan.status() == 0
    plan.close()
    # the CSR arrays of L are the CSC arrays of L^T: BACKWARD substitution on them
    indeg = oracle.csc_analyser(ci, n, nnz)
    if oracle.ref_available("d64"):
        want = oracle.ref_csc_executor(rp, ci, v, indeg, n, n, 1, 1, g)
    else:
        want = oracle.csc_executor(rp, ci, v, indeg, n, n, 1, 1, g)
    assert np.array_equal(gb, want)
    assert np.array_equal(want, oracle.adjoint(rp, ci, v, g, fused=False))


@pytest.mark.parametrize("algo", ["csc", "partition"])
def test_reference_arithmetic_refused_where_order_differs(cuda, algo):
    from capellinisptrsv_amd import CapError
    rp, ci, v = api.make_config(2, seed=0, scale=0.02)
    with pytest.raises(CapError) as e:
        _plan(cuda, rp, ci, v, algo=algo)
    assert e.value.code == -10
