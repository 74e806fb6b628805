"""GPU tests of the partitioned banded solve (CAP_ALGO_PARTITION).

The partition reassociates the sums (block solves joined by a scan over the
block tails of W = 128 values, or 256 for bandwidths 129 to 256), so its bar
is the FP64 tolerance: relative max-norm error <= 1e-12 against the oracle's
sequential forward substitution.  Cases: bandwidths 1 (one row per level),
16, 100 and 128 (the narrow tails' limit), 129, 200 and 256 (the wide
tails, the batch's deepest matrix among them), block counts
from 2 up, a ragged last block, duplicated entries, repeated solves (same
bits each time), rows longer than the 64 entries of an ELL slab, several
right-hand sides (1 to 64, interleaved; each column against its own
sequential solve, and the work vectors growing and shrinking between
solves), refinement (FP64 partitioned solves and one correction, one and
several right-hand sides), FP32 plans (FP64 tables, b widened and x rounded
once: within the FP32 tolerance, 1e-5, of the oracle's FP32 sequential
solve, and within FP32 rounding of the FP64 solve of the same FP32 system),
and the refusals (bandwidth above 256, too few rows for two blocks, blocks
whose inverses grow: the plan-creation check that compares the tables'
solve of two right-hand sides with the sequential solve).
"""
import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import api
from conftest import random_lower, rel_maxnorm

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def _chain(n):
    rp = np.r_[0, 1, np.arange(3, 2 * n, 2)].astype(np.int32)
    ci = np.r_[0, np.c_[np.arange(0, n - 1), np.arange(1, n)].reshape(-1)].astype(np.int32)
    return rp, ci, np.r_[1.0, np.tile([0.5, 2.0], n - 1)]


def _dense_band(n, w, seed):
    """Every entry within the bandwidth present: rows of up to w off-diagonals
    (past the 64 a row keeps in its ELL slab)."""
    rng = np.random.default_rng(seed)
    rp, ci, v = [0], [], []
    for i in range(n):
        cols = np.arange(max(0, i - w), i)
        off = rng.uniform(-1, 1, size=len(cols))
        ci += list(cols) + [i]
        v += list(off) + [1.0 + np.abs(off).sum()]
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(v)


def _cases():
    yield "deep_band_20k", api.make_config(2, seed=5, scale=0.04)  # bandwidth 128, ~38 nnz/row
    yield "chain_5000", _chain(5000)
    yield "band16", api.generate("band_rowcount", n=9000, offdiag=6, band=16, seed=3)
    yield "band100_ragged", api.generate("band_rowcount", n=70001, offdiag=12, band=100, seed=4)
    yield "band128_two_blocks", api.generate("band_rowcount", n=200, offdiag=10, band=128, seed=6)
    yield "dup", random_lower(4000, max_off=6, band=90, seed=7, dup=True)
    yield "dense_band128", _dense_band(3000, 128, seed=8)
    yield "band129", api.generate("band_rowcount", n=40000, offdiag=9, band=129, seed=10)
    yield "band200_ragged", api.generate("band_rowcount", n=90001, offdiag=12, band=200, seed=11)
    yield "band256_two_blocks", api.generate("band_rowcount", n=600, offdiag=10, band=256, seed=12)
    yield "dense_band256", _dense_band(3000, 256, seed=13)
    yield "batch_deepest", _batch_deepest()


def _batch_deepest():
    """The deepest matrix of the 256-matrix batch workload: n = 158,094,
    16,132 levels, lower bandwidth 129."""
    mats = api.make_batch(256, seed=0)[0]
    depth = [api.matrix_layer(len(rp) - 1, len(rp) - 1, len(ci), rp, ci, v)[0] for rp, ci, v in mats]
    best = int(np.argmax(depth))
    return mats[best]


CASES = list(_cases())
IDX = {c[0]: i for i, c in enumerate(CASES)}


def _plan(cuda, rp, ci, v, **kw):
    import torch
    from capellinisptrsv_amd import Plan
    return Plan(*(torch.from_numpy(a).to(cuda) for a in (rp, ci, v)), **kw)


@pytest.mark.parametrize("name,mat", CASES, ids=[c[0] for c in CASES])
def test_partition_within_fp64_tolerance(cuda, name, mat):
    import torch
    rp, ci, v = mat
    n = len(rp) - 1
    b = api.uniform(31, n)
    plan = _plan(cuda, rp, ci, v, algo="partition")
    assert plan.info()["algo"] == "partition"
    bt = torch.from_numpy(b).to(cuda)
    x1 = plan.solve(bt).cpu().numpy()
    x2 = plan.solve(bt).cpu().numpy()
    assert plan.status() == 0
    ref = oracle.csr_forward(rp, ci, v, b, fused=True)
    assert rel_maxnorm(x1, ref) <= FP64_TOL, rel_maxnorm(x1, ref)
    assert np.array_equal(x1, x2)  # deterministic: same bits on every solve
    plan.close()


def _columns_within_tol(x, ref, k):
    x, ref = x.reshape(-1, k), ref.reshape(-1, k)
    errs = [rel_maxnorm(x[:, q], ref[:, q]) for q in range(k)]
    assert max(errs) <= FP64_TOL, errs


SEVERAL = ("deep_band_20k", "band100_ragged", "dense_band128", "band200_ragged", "band256_two_blocks")


@pytest.mark.parametrize("name,mat", [CASES[IDX[c]] for c in SEVERAL], ids=SEVERAL)
def test_partition_several_rhs(cuda, name, mat):
    """k = 16, 64, then smaller counts (the work vectors grow, then serve a
    smaller k as they are), with chunks of 16, 8, 4, 2 and 1 right-hand sides
    per workgroup (k = 12: 4; k = 6: 2; k = 3: 1).  Each column is within the
    tolerance of its own sequential solve and has the bits of the same column
    solved alone (the chunked kernels keep the one-RHS summation order)."""
    import torch
    rp, ci, v = mat
    n = len(rp) - 1
    plan = _plan(cuda, rp, ci, v, algo="partition")

    def alone(col):
        return plan.solve(torch.from_numpy(np.ascontiguousarray(col)).to(cuda)).cpu().numpy()

    for seed, k in ((40, 16), (41, 64), (42, 2), (43, 1), (44, 4), (45, 8), (46, 12), (47, 6), (48, 3)):
        b = api.uniform(seed, n * k).reshape(n, k)
        x = plan.solve(torch.from_numpy(b).to(cuda)).cpu().numpy().reshape(n, k)
        assert plan.status() == 0
        _columns_within_tol(x, oracle.csr_forward(rp, ci, v, b.reshape(-1), rhs=k, fused=True), k)
        for q in sorted({0, k // 2, k - 1}):
            assert np.array_equal(x[:, q], alone(b[:, q])), (k, q)
        if k == 1:
            # the k = 1 solve through work vectors sized for 64 gives the bits of a fresh plan
            fresh = _plan(cuda, rp, ci, v, algo="partition")
            assert np.array_equal(x[:, 0], fresh.solve(torch.from_numpy(b[:, 0].copy()).to(cuda)).cpu().numpy())
            fresh.close()
    plan.close()


@pytest.mark.parametrize("k", [1, 4])
def test_partition_refinement(cuda, k):
    import torch
    rp, ci, v = api.make_config(2, seed=8, scale=0.02)
    n = len(rp) - 1
    plan = _plan(cuda, rp, ci, v, algo="partition")
    b = api.uniform(33, n * k)
    bt = torch.from_numpy(b.reshape(n, k) if k > 1 else b).to(cuda)
    xr, iters, rel = plan.solve_refined(bt, tol=1e-12)
    # FP64 partitioned solves: the first correction is already below 1e-12
    assert iters == 2 and rel <= 1e-12, (iters, rel)
    _columns_within_tol(xr.cpu().numpy().reshape(-1), oracle.csr_forward(rp, ci, v, b, rhs=k, fused=True), k)
    plan.close()


FP32_CASES = ("deep_band_20k", "band100_ragged", "dense_band128", "band129")


@pytest.mark.parametrize("name,mat", [CASES[IDX[c]] for c in FP32_CASES], ids=FP32_CASES)
@pytest.mark.parametrize("k", [1, 4])
def test_partition_fp32(cuda, name, mat, k):
    import torch
    rp, ci, v = mat
    n = len(rp) - 1
    v32 = v.astype(np.float32)
    b32 = api.uniform(35, n * k).astype(np.float32)
    plan = _plan(cuda, rp, ci, v32, algo="partition")
    assert plan.info()["algo"] == "partition"
    x = plan.solve(torch.from_numpy(b32.reshape(n, k) if k > 1 else b32).to(cuda)).cpu().numpy().reshape(n, k)
    assert plan.status() == 0 and x.dtype == np.float32
    ref32 = oracle.csr_forward(rp, ci, v32, b32, rhs=k, fused=True).reshape(n, k)
    # the same FP32 system solved in FP64: the partition is that, rounded once
    ref64 = oracle.csr_forward(rp, ci, v32.astype(np.float64), b32.astype(np.float64), rhs=k, fused=True).reshape(n, k)
    for q in range(k):
        assert rel_maxnorm(x[:, q], ref32[:, q]) <= FP32_TOL, (q, rel_maxnorm(x[:, q], ref32[:, q]))
        assert rel_maxnorm(x[:, q].astype(np.float64), ref64[:, q]) <= 1e-7, q
    plan.close()


def test_partition_refusals(cuda):
    from capellinisptrsv_amd import CapError
    rp, ci, v = api.generate("band_rowcount", n=5000, offdiag=8, band=300, seed=9)
    with pytest.raises(CapError):
        _plan(cuda, rp, ci, v, algo="partition")  # bandwidth 300 > 256
    rp, ci, v = _dense_band(1500, 257, seed=14)
    with pytest.raises(CapError):
        _plan(cuda, rp, ci, v, algo="partition")  # bandwidth 257 > 256
    rp, ci, v = api.generate("band_rowcount", n=250, offdiag=8, band=200, seed=9)
    with pytest.raises(CapError):
        _plan(cuda, rp, ci, v, algo="partition")  # one block at the wide tails' 256 rows
    rp, ci, v = api.generate("band_rowcount", n=128, offdiag=4, band=64, seed=9)
    with pytest.raises(CapError):
        _plan(cuda, rp, ci, v, algo="partition")  # one block


def _unstable(n):
    """Unit values in columns i-1 and i-3 and on the diagonal: the recurrence
    x_i = -(x_{i-1} + x_{i-3}) has a root of modulus 1.47, so the block
    inverses grow like 1.47^m while the sequential solve of L x = L 1 stays
    exact."""
    rp, ci = [0], []
    for i in range(n):
        ci += [c for c in (i - 3, i - 1) if c >= 0] + [i]
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.ones(len(ci))


def test_partition_refuses_growing_blocks_and_auto_avoids_it(cuda):
    import torch
    from capellinisptrsv_amd import CapError
    rp, ci, v = _unstable(3000)
    with pytest.raises(CapError):
        _plan(cuda, rp, ci, v, algo="partition")
    plan = _plan(cuda, rp, ci, v)  # autotuned: the partition is no candidate
    assert plan.info()["algo"] != "partition"
    b = api.uniform(36, len(rp) - 1)
    x = plan.solve(torch.from_numpy(b).to(cuda)).cpu().numpy()
    assert plan.status() == 0
    assert np.array_equal(x, oracle.csr_forward(rp, ci, v, b, fused=True))
    plan.close()


def _growing_bidiagonal(n, a=33.0 / 32.0):
    """x_i = b_i + a x_{i-1}: unit diagonal, sub-diagonal -a.  With a = 33/32
    and b = L 1 every step of the sequential solve is exact (x = 1), while the
    block inverses hold a^k (a^127 = 50) and the scan over the block tails
    compounds them: on four blocks the partitioned solve is off by about
    1e-11 -- past the FP64 bar of 1e-12, yet inside the 1e-10 absolute test
    on L x = L 1 the plan-creation check used to apply."""
    rp = np.r_[0, np.arange(1, 2 * n, 2)].astype(np.int32)
    ci = np.r_[0, np.c_[np.arange(0, n - 1), np.arange(1, n)].reshape(-1)].astype(np.int32)
    return rp, ci, np.r_[1.0, np.tile([-a, 1.0], n - 1)]


@pytest.mark.parametrize("n", [512, 640])
def test_partition_refuses_moderate_growth_and_auto_avoids_it(cuda, n):
    """The plan-creation check holds the partition to 1e-13 relative of the
    sequential solve (two right-hand sides: L 1 and a seeded U[-1,1] b)."""
    import torch
    from capellinisptrsv_amd import CapError
    rp, ci, v = _growing_bidiagonal(n)
    with pytest.raises(CapError) as e:
        _plan(cuda, rp, ci, v, algo="partition")
    assert e.value.code == -10  # CAP_ERR_UNSUPPORTED
    plan = _plan(cuda, rp, ci, v)
    assert plan.info()["algo"] != "partition"
    b = np.asarray([v[rp[i]:rp[i + 1]].sum() for i in range(n)])  # L 1: the sequential solve gives x = 1
    x = plan.solve(torch.from_numpy(b).to(cuda)).cpu().numpy()
    assert plan.status() == 0
    assert np.array_equal(x, oracle.csr_forward(rp, ci, v, b, fused=True))
    assert np.array_equal(x, np.ones(n))
    plan.close()


def test_autotuner_on_a_deep_band(cuda):
    import torch
    rp, ci, v = api.make_config(2, seed=10, scale=0.04)
    n = len(rp) - 1
    plan = _plan(cuda, rp, ci, v)
    info = plan.info()
    b = api.uniform(34, n)
    x = plan.solve(torch.from_numpy(b).to(cuda)).cpu().numpy()
    assert plan.status() == 0
    ref = oracle.csr_forward(rp, ci, v, b, fused=True)
    if info["algo"] == "partition":
        assert rel_maxnorm(x, ref) <= FP64_TOL
    else:
        assert np.array_equal(x, ref), info
    plan.close()
