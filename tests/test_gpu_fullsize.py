"""GPU parity at the benchmark workloads' full sizes, through the C ABI.

Each case runs the HIP path on the workload exactly as the benchmark builds it
and checks every value against the oracle:

  * wide-level n = 2,000,000 (200 levels x 10,000 rows), FP64: the autotuned
    plan and the write-first kernel bit-exact against the oracle's fused
    sequential forward substitution, and within 1e-12 relative max-norm of
    the reference's own serial executor (openCL/SynFree_csc/
    sptrsv_syncfree_serialref.h:44-72, built from its sources into
    oracle/_ref; the oracle's restatement of it, bit-equal to it on the CPU,
    where that build is absent).  FP32: within 1e-5 of the FP64 oracle.
  * deep band n = 500,000 (bandwidth 128, p = 0.3), FP64: the autotuned plan
    (the partitioned solve, reassociated sums) within 1e-12; the mix kernel
    (every row a wavefront row at ~39 nnz/row) and the one-workgroup kernel
    bit-exact.  FP32 autotuned within 1e-5 of the FP64 oracle.
  * the batch: 256 matrices, n log-uniform in [2,000, 200,000], 16 RHS each:
    the FP64 batch solve bit-exact for every matrix and every RHS, and the
    FP32 solve inside FP64 refinement within 1e-12 of each matrix's FP64
    oracle solve.
"""
import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import api
from conftest import rel_maxnorm

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def _serial_reference(rp, ci, v, b):
    """The reference's column-driven serial executor on the same system."""
    n, nnz = len(rp) - 1, len(ci)
    cr, cp, cv = oracle.transpose(n, n, nnz, rp, ci, v)
    indeg = oracle.csc_analyser(cr, n, nnz)
    if oracle.ref_available("d64"):
        return oracle.ref_csc_executor(cp, cr, cv, indeg, n, n, 0, 1, b)
    return oracle.csc_executor(cp, cr, cv, indeg, n, n, 0, 1, b)


@pytest.fixture(scope="module")
def wide_level():
    rp, ci, v = api.make_config(1, seed=0)
    b = api.uniform(1, len(rp) - 1)
    return rp, ci, v, b, oracle.csr_forward(rp, ci, v, b, fused=True)


@pytest.fixture(scope="module")
def deep_band():
    rp, ci, v = api.make_config(2, seed=0)
    b = api.uniform(3, len(rp) - 1)
    return rp, ci, v, b, oracle.csr_forward(rp, ci, v, b, fused=True)


def _plan_solve(cuda, rp, ci, v, b, **kw):
    import torch
    from capellinisptrsv_amd import Plan
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v)]
    plan = Plan(*d, **kw)
    x = plan.solve(torch.from_numpy(b.astype(v.dtype)).to(cuda)).cpu().numpy()
    assert plan.status() == 0
    info = plan.info()
    plan.close()
    return x, info


@pytest.mark.parametrize("algo", ["auto", "wrtfst"])
def test_wide_level_2m_fp64(cuda, wide_level, algo):
    rp, ci, v, b, ref = wide_level
    assert len(rp) - 1 == 2_000_000
    x, info = _plan_solve(cuda, rp, ci, v, b, algo=algo)
    assert info["algo"] in ("wrtfst", "mix", "twopart"), info
    assert np.array_equal(x.view(np.uint64), ref.view(np.uint64)), info
    assert rel_maxnorm(x, _serial_reference(rp, ci, v, b)) <= FP64_TOL


def test_wide_level_2m_fp32(cuda, wide_level):
    rp, ci, v, b, ref = wide_level
    v32 = v.astype(np.float32)
    x, info = _plan_solve(cuda, rp, ci, v32, b.astype(np.float32), algo="auto")
    assert rel_maxnorm(x, ref) <= FP32_TOL, info
    # the FP32 kernels keep the sequential order: bit-exact against the FP32 oracle
    assert np.array_equal(x, oracle.csr_forward(rp, ci, v32, b.astype(np.float32), fused=True)), info


def test_deep_band_500k_autotuned(cuda, deep_band):
    rp, ci, v, b, ref = deep_band
    assert len(rp) - 1 == 500_000
    x, info = _plan_solve(cuda, rp, ci, v, b, algo="auto")
    if info["algo"] == "partition":  # reassociated sums: the FP64 tolerance
        assert rel_maxnorm(x, ref) <= FP64_TOL, info
    else:
        assert np.array_equal(x, ref), info


@pytest.mark.parametrize("algo", ["mix", "local"])
def test_deep_band_500k_row_kernels_bit_exact(cuda, deep_band, algo):
    rp, ci, v, b, ref = deep_band
    x, info = _plan_solve(cuda, rp, ci, v, b, algo=algo)
    if algo == "mix":
        assert info["warp_occupy"] > 0.99, info  # wavefront rows throughout
    assert np.array_equal(x.view(np.uint64), ref.view(np.uint64)), info


def test_deep_band_500k_fp32(cuda, deep_band):
    rp, ci, v, b, ref = deep_band
    x, info = _plan_solve(cuda, rp, ci, v.astype(np.float32), b.astype(np.float32), algo="auto")
    assert rel_maxnorm(x, ref) <= FP32_TOL, info


@pytest.fixture(scope="module")
def batch256():
    mats, row_off, rp, ci, v = api.make_batch(256, seed=0)
    return mats, row_off, rp, ci, v


def test_batch_256x16_fp64_and_refined(cuda, batch256):
    import torch
    from capellinisptrsv_amd import Batch
    mats, row_off, rp, ci, v = batch256
    k = 16
    N = int(row_off[-1])
    b = api.uniform(7, N * k)
    t = [torch.from_numpy(a).to(cuda) for a in (row_off, rp, ci, v)]
    bt = Batch(*t, k=k, algo="auto")
    bd = torch.from_numpy(b.reshape(N, k)).to(cuda)
    x = bt.solve(bd).cpu().numpy().reshape(-1)
    xr, iters, rel = bt.solve_refined(bd, tol=1e-12)
    xr = xr.cpu().numpy().reshape(-1)
    part = set(bt.partitioned_matrices)  # the hybrid batch's partitioned matrices: within 1e-12
    bt.close()
    for m, (mrp, mci, mv) in enumerate(mats):
        lo, hi = row_off[m] * k, row_off[m + 1] * k
        ref = oracle.csr_forward(mrp, mci, mv, b[lo:hi], rhs=k, fused=True)
        if m in part:
            assert rel_maxnorm(x[lo:hi], ref) <= FP64_TOL, m
        else:
            assert np.array_equal(x[lo:hi], ref), m
        assert rel_maxnorm(xr[lo:hi], ref) <= FP64_TOL, (m, iters, rel)
