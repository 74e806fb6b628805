"""GPU FP32-solve + FP64-refinement and batched multi-matrix / multi-RHS solves.

Refinement bar: relative max-norm error <= 1e-12 against the FP64 oracle
(the FP32-solve + FP64-refinement accuracy target).  Batched bar: each matrix's k
right-hand sides bit-exact against the oracle's multi-RHS forward
substitution on that matrix (the same sequential fused order).
"""
import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import api
from conftest import rel_maxnorm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,scale,k,algo", [(1, 0.05, 1, "wrtfst"), (0, 1.0, 1, "wrtfst"), (0, 0.2, 16, "wrtfst"),
                                              (4, 0.1, 8, "wrtfst"), (2, 0.02, 1, "local"), (0, 0.2, 16, "local")])
def test_refined_reaches_1e12(cuda, cfg, scale, k, algo):
    import torch
    from capellinisptrsv_amd import Plan
    rp, ci, v = api.make_config(cfg, seed=1, scale=scale)
    n = len(rp) - 1
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v)]
    plan = Plan(*d, algo=algo, sync="value")
    b = api.uniform(9, n * k)
    bt = torch.from_numpy(b.reshape(n, k) if k > 1 else b).to(cuda)
    out = torch.empty_like(bt)
    for rep in range(2):  # same b and x: the second call replays the captured graph
        x, iters, rel = plan.solve_refined(bt, tol=1e-12, max_iter=10, out=out)
        torch.cuda.synchronize()
        ref = oracle.csr_forward(rp, ci, v, b, rhs=k, fused=True)
        err = rel_maxnorm(x.cpu().numpy().reshape(-1), ref)
        assert err <= 1e-12, (rep, iters, rel, err)
        assert 2 <= iters <= 4


@pytest.mark.parametrize("algo,readiness,lds,bpc,k,threads", [
    ("wrtfst", "staged", True, 1, 1, 0), ("wrtfst", "staged", False, 2, 1, 0), ("wrtfst", "in_place", True, 1, 1, 0),
    ("wrtfst", "in_place", False, 4, 1, 0), ("wrtfst", "staged", True, 1, 8, 0),
    ("wrtfst", "in_place", False, 2, 16, 0), ("wrtfst", "staged", True, 1, 1, 128),
    ("wrtfst", "in_place", False, 1, 16, 128), ("local", "auto", False, 1, 1, 0), ("local", "auto", False, 1, 16, 0)])
def test_refine_two_steps_bit_exact(cuda, algo, readiness, lds, bpc, k, threads):
    """Thread-per-row plans run the first two refinement steps as one fused
    sweep (FP32 rows end with a product by the reciprocal diagonal),
    one-workgroup plans as two solves with the residual between them
    (divisions); both must give the oracle's two-step result in that form
    bit for bit."""
    import torch
    from capellinisptrsv_amd import Plan
    for name, (rp, ci, v) in (("cfg1_small", api.make_config(1, seed=2, scale=0.05)),
                              ("band", api.generate("band_rowcount", n=6000, offdiag=12, band=300, seed=5))):
        n = len(rp) - 1
        d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v)]
        plan = Plan(*d, algo=algo, sync="value", lds_window=lds, blocks_per_cu=bpc, readiness=readiness,
                    row_threads=threads)
        b = api.uniform(21, n * k)
        bt = torch.from_numpy(b.reshape(n, k) if k > 1 else b).to(cuda)
        _, _, want = oracle.refine2(rp, ci, v, b, rhs=k, rcp=algo != "local")  # the sweep's reciprocal form
        out = torch.empty_like(bt)
        for rep in range(2):  # same b and x: the second call replays the captured graph
            x, iters, rel = plan.solve_refined(bt, tol=1e-12, max_iter=10, out=out)
            torch.cuda.synchronize()
            assert plan.status() == 0
            assert iters == 2, (name, rep, iters, rel)
            assert np.array_equal(x.cpu().numpy().reshape(-1), want), (name, rep)
        plan.close()


def _batch(seed, n_mat, k, nmin=200, nmax=5000):
    rng = np.random.default_rng(seed)
    mats = []
    for m in range(n_mat):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        band = int(rng.integers(64, 4097))
        mats.append(api.generate("band_rowcount", n=n, offdiag=16, band=band, seed=seed * 1000 + m))
    row_off = np.cumsum([0] + [len(rp) - 1 for rp, _, _ in mats]).astype(np.int32)
    nnz_off = np.cumsum([0] + [len(ci) for _, ci, _ in mats])
    rp_all = np.concatenate([rp[:-1] + o for (rp, _, _), o in zip(mats, nnz_off)] + [[nnz_off[-1]]]).astype(np.int32)
    ci_all = np.concatenate([ci for _, ci, _ in mats]).astype(np.int32)
    v_all = np.concatenate([v for _, _, v in mats])
    return mats, row_off, rp_all, ci_all, v_all


@pytest.mark.parametrize("algo", ["wrtfst", "local", "auto"])
@pytest.mark.parametrize("k", [1, 8, 16, 32])
def test_batch_bit_exact_per_matrix(cuda, k, algo):
    import torch
    from capellinisptrsv_amd import Batch
    mats, row_off, rp, ci, v = _batch(3, 24, k)
    N = row_off[-1]
    b = api.uniform(11, N * k)
    t = [torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v)]
    bt = Batch(*t, k=k, algo=algo)
    if algo != "auto":
        assert bt.algo == algo
    x = bt.solve(torch.from_numpy(b.reshape(N, k)).to(cuda)).cpu().numpy().reshape(-1)
    part = set(bt.partitioned_matrices)  # a hybrid batch's partitioned matrices: within 1e-12
    for m, (mrp, mci, mv) in enumerate(mats):
        lo, hi = row_off[m] * k, row_off[m + 1] * k
        ref = oracle.csr_forward(mrp, mci, mv, b[lo:hi], rhs=k, fused=True)
        if m in part:
            assert rel_maxnorm(x[lo:hi], ref) <= 1e-12, m
        else:
            assert np.array_equal(x[lo:hi], ref), m


def _hybrid_batch():
    """Deep banded matrices (bandwidths 3, 129 and 200, thousands of levels)
    among 20 shallow wide ones: the deep chains dominate the batch."""
    mats = [api.generate("band", n=60000, offdiag=2, band=3, seed=81),
            api.generate("band_rowcount", n=40000, offdiag=9, band=129, seed=82),
            api.generate("band_rowcount", n=30000, offdiag=9, band=200, seed=83)]
    mats += [api.generate("level", n=8000, levels=4, width=2000, offdiag=6, depth=2, seed=90 + i) for i in range(20)]
    return _concat(mats)


def _concat(mats):
    row_off = np.cumsum([0] + [len(m[0]) - 1 for m in mats]).astype(np.int32)
    nnz_off = np.cumsum([0] + [len(m[1]) for m in mats])
    rp = np.concatenate([m[0][:-1] + o for m, o in zip(mats, nnz_off)] + [[nnz_off[-1]]]).astype(np.int32)
    return mats, row_off, rp, np.concatenate([m[1] for m in mats]).astype(np.int32), np.concatenate(
        [m[2] for m in mats])


def _growing(n=3000, lag=200, c=1.9):
    """x_i = b_i - c x_{i-lag}: 15 levels.  With b = L x for an x of order 1,
    each level multiplies the error of the one before by 1.9, so an FP32
    solve is accurate to ~1e-3 and refinement needs several steps."""
    rp, ci, v = [0], [], []
    for i in range(n):
        if i >= lag:
            ci.append(i - lag)
            v.append(c)
        ci.append(i)
        v.append(1.0)
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(v)


def test_hybrid_batch_refined_more_steps(cuda):
    """A hybrid batch whose thread-per-row part needs more than two
    refinement steps: the later steps (residual, FP32 correction, x += d over
    the batch) must leave the partitioned matrices' rows as their FP64
    solves left them, and every matrix ends within 1e-12."""
    import torch
    from capellinisptrsv_amd import Batch
    mats = _hybrid_batch()[0] + [_growing()]
    mats, row_off, rp, ci, v = _concat(mats)
    k = 4
    N = int(row_off[-1])
    b = api.uniform(15, N * k)
    grp, gci, gv = mats[-1]
    rows = np.repeat(np.arange(len(grp) - 1), np.diff(grp))
    xt = api.uniform(16, (len(grp) - 1) * k).reshape(-1, k)
    bg = np.zeros_like(xt)
    np.add.at(bg, rows, gv[:, None] * xt[gci])
    b[row_off[-2] * k:] = bg.reshape(-1)
    bt = Batch(*(torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v)), k=k, algo="auto")
    assert bt.partitioned >= 1
    bd = torch.from_numpy(b.reshape(N, k)).to(cuda)
    x = bt.solve(bd).cpu().numpy().reshape(-1)
    xr, iters, rel = bt.solve_refined(bd, tol=1e-13, max_iter=10)
    xr = xr.cpu().numpy().reshape(-1)
    assert iters >= 3, (iters, rel)
    part = set(bt.partitioned_matrices)
    for m, (mrp, mci, mv) in enumerate(mats):
        lo, hi = row_off[m] * k, row_off[m + 1] * k
        ref = oracle.csr_forward(mrp, mci, mv, b[lo:hi], rhs=k, fused=True)
        if m in part:
            assert np.array_equal(xr[lo:hi], x[lo:hi]), m
        # the growing member's condition number (~1.9^15) leaves the FP64
        # oracle itself ~1e-12 from the exact solution
        assert rel_maxnorm(xr[lo:hi], ref) <= (1e-10 if m == len(mats) - 1 else 1e-12), (m, iters, rel)
    bt.close()


@pytest.mark.parametrize("k", [1, 16])
def test_hybrid_batch(cuda, k):
    """An autotuned batch whose deep matrices are solved by partitioned plans
    of their own: those within 1e-12 of their sequential solves, the others
    bit-exact; the refined solve within 1e-12 everywhere; repeated solves give
    the same bits."""
    import torch
    from capellinisptrsv_amd import Batch
    mats, row_off, rp, ci, v = _hybrid_batch()
    N = int(row_off[-1])
    b = api.uniform(13, N * k)
    t = [torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v)]
    bt = Batch(*t, k=k, algo="auto")
    assert bt.algo == "wrtfst" and bt.partitioned >= 1 and set(bt.partitioned_matrices) <= {0, 1, 2}
    # row tiles only where the rest runs 1024-thread workgroups, one right-hand side per lane
    assert not bt.row_tiles or (bt.row_threads == 1024 and bt.rhs_per_lane == 1)
    bd = torch.from_numpy(b.reshape(N, k)).to(cuda)
    _check_hybrid(bt, bd, mats, row_off, b, k)
    bt.close()


def test_hybrid_batch_fp32(cuda):
    """The FP32 hybrid batch: FP32 partitioned plans (FP64 tables, b widened
    and x rounded once) within 1e-5 of the FP64 solve of the same FP32
    system, the thread-per-row matrices bit-exact against the FP32 oracle."""
    import torch
    from capellinisptrsv_amd import Batch
    mats, row_off, rp, ci, v = _hybrid_batch()
    k = 4
    N = int(row_off[-1])
    b = api.uniform(14, N * k).astype(np.float32)
    t = [torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v.astype(np.float32))]
    bt = Batch(*t, k=k, algo="auto")
    assert bt.partitioned >= 1
    x = bt.solve(torch.from_numpy(b.reshape(N, k)).to(cuda)).cpu().numpy().reshape(-1)
    part = set(bt.partitioned_matrices)
    for m, (mrp, mci, mv) in enumerate(mats):
        lo, hi = row_off[m] * k, row_off[m + 1] * k
        v32 = mv.astype(np.float32)
        if m in part:
            ref64 = oracle.csr_forward(mrp, mci, v32.astype(np.float64), b[lo:hi].astype(np.float64), rhs=k)
            assert rel_maxnorm(x[lo:hi], ref64) <= 1e-5, m
        else:
            assert np.array_equal(x[lo:hi], oracle.csr_forward(mrp, mci, v32, b[lo:hi], rhs=k, fused=True)), m
    bt.close()


def _check_hybrid(bt, bd, mats, row_off, b, k):
    import torch
    x = bt.solve(bd).cpu().numpy().reshape(-1)
    x2 = bt.solve(bd).cpu().numpy().reshape(-1)
    assert np.array_equal(x, x2)
    xr, iters, rel = bt.solve_refined(bd, tol=1e-12)
    xr = xr.cpu().numpy().reshape(-1)
    # below the partitioned plans' creation check (1e-13) they are refined in
    # FP64 too; a correction that stalls near the rounding noise of an
    # ill-conditioned member may report non-convergence, the result must hold
    from capellinisptrsv_amd import CapError
    xt = torch.empty_like(bd)
    iters_t = rel_t = None
    try:
        _, iters_t, rel_t = bt.solve_refined(bd, tol=9e-14, out=xt)
    except CapError as e:
        assert e.code == -12, e
    xt = xt.cpu().numpy().reshape(-1)
    part = set(bt.partitioned_matrices)
    for m, (mrp, mci, mv) in enumerate(mats):
        lo, hi = row_off[m] * k, row_off[m + 1] * k
        ref = oracle.csr_forward(mrp, mci, mv, b[lo:hi], rhs=k, fused=True)
        if m in part:
            assert rel_maxnorm(x[lo:hi], ref) <= 1e-12, m
            # at tol 1e-12 the refined batch solves them once in FP64
            assert np.array_equal(xr[lo:hi], x[lo:hi]), m
        else:
            assert np.array_equal(x[lo:hi], ref), m
        assert rel_maxnorm(xr[lo:hi], ref) <= 1e-12, (m, iters, rel)
        assert rel_maxnorm(xt[lo:hi], ref) <= 1e-12, (m, iters_t, rel_t)


def _chain(n):
    """Bidiagonal matrix: one row per level (depth n)."""
    rp = np.r_[0, 1, np.arange(3, 2 * n, 2)].astype(np.int32)
    ci = np.r_[0, np.c_[np.arange(0, n - 1), np.arange(1, n)].reshape(-1)].astype(np.int32)
    return rp, ci, np.r_[1.0, np.tile([0.5, 2.0], n - 1)]


@pytest.mark.parametrize("bpc", [1, 2, 4])
@pytest.mark.parametrize("k", [1, 16])
def test_batch_grid_sizes_with_mixed_depths(cuda, k, bpc):
    # a chain of depth 3000 next to wide band matrices: the schedule orders
    # blocks by level / depth, so the chain's blocks are interleaved early
    import torch
    from capellinisptrsv_amd import Batch
    mats, _, _, _, _ = _batch(7, 12, k)
    mats.insert(5, _chain(3000))
    row_off = np.cumsum([0] + [len(m[0]) - 1 for m in mats]).astype(np.int32)
    nnz_off = np.cumsum([0] + [len(m[1]) for m in mats])
    rp = np.concatenate([m[0][:-1] + o for m, o in zip(mats, nnz_off)] + [[nnz_off[-1]]]).astype(np.int32)
    ci = np.concatenate([m[1] for m in mats]).astype(np.int32)
    v = np.concatenate([m[2] for m in mats])
    N = row_off[-1]
    b = api.uniform(13, N * k)
    t = [torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v)]
    bt = Batch(*t, k=k, algo="wrtfst", blocks_per_cu=bpc)
    assert (bt.algo, bt.blocks_per_cu) == ("wrtfst", bpc)
    x = bt.solve(torch.from_numpy(b.reshape(N, k) if k > 1 else b).to(cuda)).cpu().numpy().reshape(-1)
    for m, (mrp, mci, mv) in enumerate(mats):
        lo, hi = row_off[m] * k, row_off[m + 1] * k
        assert np.array_equal(x[lo:hi], oracle.csr_forward(mrp, mci, mv, b[lo:hi], rhs=k, fused=True)), m
    bt.close()


def test_batch_refuses_more_than_four_blocks_per_cu(cuda):
    import torch
    from capellinisptrsv_amd import Batch
    _, row_off, rp, ci, v = _batch(3, 4, 1)
    t = [torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v)]
    with pytest.raises(RuntimeError, match="cap_batch_create"):
        Batch(*t, k=1, algo="wrtfst", blocks_per_cu=8)


@pytest.mark.parametrize("algo", ["wrtfst", "local"])
def test_batch_refined(cuda, algo):
    import torch
    from capellinisptrsv_amd import Batch
    k = 16
    mats, row_off, rp, ci, v = _batch(5, 16, k)
    N = row_off[-1]
    b = api.uniform(12, N * k)
    t = [torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v)]
    bt = Batch(*t, k=k, algo=algo)
    x, iters, rel = bt.solve_refined(torch.from_numpy(b.reshape(N, k)).to(cuda), tol=1e-12)
    x = x.cpu().numpy().reshape(-1)
    for m, (mrp, mci, mv) in enumerate(mats):
        lo, hi = row_off[m] * k, row_off[m + 1] * k
        ref = oracle.csr_forward(mrp, mci, mv, b[lo:hi], rhs=k, fused=True)
        assert rel_maxnorm(x[lo:hi], ref) <= 1e-12, (m, iters, rel)
        if iters == 2:  # stopped after the first correction: the oracle's two steps exactly
            want = oracle.refine2(mrp, mci, mv, b[lo:hi], rhs=k, rcp=algo != "local")[2]
            assert np.array_equal(x[lo:hi], want), m


def test_batch_rejects_bad_matrix(cuda):
    import torch
    from capellinisptrsv_amd import Batch, CapError
    mats, row_off, rp, ci, v = _batch(6, 4, 1)
    ci = ci.copy()
    ci[rp[row_off[2] + 5 + 1] - 1] = 0  # a diagonal of matrix 2 replaced: no diagonal
    t = [torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v)]
    with pytest.raises(CapError):
        Batch(*t, k=1)


@pytest.mark.parametrize("algo", ["wrtfst", "local"])
def test_batch_with_empty_and_single_row_matrices(cuda, algo):
    """Empty matrices (n = 0) and 1-row matrices inside a batch."""
    import torch
    from capellinisptrsv_amd import Batch
    mats = [api.generate("band_rowcount", n=n, offdiag=8, band=64, seed=70 + i) if n else
            (np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
            for i, n in enumerate([300, 0, 1, 0, 2000, 1])]
    row_off = np.cumsum([0] + [len(m[0]) - 1 for m in mats]).astype(np.int32)
    nnz_off = np.cumsum([0] + [len(m[1]) for m in mats])
    rp = np.concatenate([m[0][:-1] + o for m, o in zip(mats, nnz_off)] + [[nnz_off[-1]]]).astype(np.int32)
    ci = np.concatenate([m[1] for m in mats]).astype(np.int32)
    v = np.concatenate([m[2] for m in mats])
    k = 2
    N = int(row_off[-1])
    b = api.uniform(5, N * k)
    t = [torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v)]
    bt = Batch(*t, k=k, algo=algo)
    x = bt.solve(torch.from_numpy(b.reshape(N, k)).to(cuda)).cpu().numpy().reshape(-1)
    for m, (mrp, mci, mv) in enumerate(mats):
        lo, hi = row_off[m] * k, row_off[m + 1] * k
        if hi > lo:
            assert np.array_equal(x[lo:hi], oracle.csr_forward(mrp, mci, mv, b[lo:hi], rhs=k, fused=True)), m
