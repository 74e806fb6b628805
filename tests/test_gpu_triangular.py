"""GPU parity of upper-triangular and implicit-unit-diagonal solves.

Plans convert such matrices into the lower, diagonal-last layout every
kernel solves (row and column reversal for upper) and reverse b and x
around the solve.  Bar: bit-exact against the oracle's stored-order
triangular solve wherever the accumulation order is the stored one (all
thread/wavefront kernels); CSC and the partitioned solve (reassociated
sums) within 1e-12.
"""
import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import api
from conftest import drop_diagonal, random_lower, rel_maxnorm, upper_from_lower

pytestmark = pytest.mark.gpu


def _mat(kind, n=3000, seed=4):
    rp, ci, v = random_lower(n, max_off=6, band=200, seed=seed, dup=True)
    upper = kind in ("upper", "upper_unit")
    unit = kind in ("unit", "upper_unit")
    if upper:
        rp, ci, v = upper_from_lower(rp, ci, v)
    if unit:
        rp, ci, v = drop_diagonal(rp, ci, v, upper=upper)
    return rp, ci, v, upper, unit


def _plan(cuda, rp, ci, v, **kw):
    import torch
    from capellinisptrsv_amd import Plan
    return Plan(*(torch.from_numpy(a).to(cuda) for a in (rp, ci, v)), **kw)


@pytest.mark.parametrize("algo", ["wrtfst", "twopart", "mix", "local", "csc", "auto"])
@pytest.mark.parametrize("kind", ["upper", "unit", "upper_unit"])
def test_triangular_kinds_match_the_oracle(cuda, kind, algo):
    import torch
    rp, ci, v, upper, unit = _mat(kind)
    n = len(rp) - 1
    b = api.uniform(41, n)
    plan = _plan(cuda, rp, ci, v, algo=algo, upper=upper, unit_diagonal=unit, border=4)
    x = plan.solve(torch.from_numpy(b).to(cuda))
    torch.cuda.synchronize()
    assert plan.status() == 0
    x = x.cpu().numpy()
    ref = oracle.csr_solve(rp, ci, v, b, upper=upper, unit=unit, fused=True)
    if algo == "csc" or (algo == "auto" and plan.info()["algo"] in ("csc", "partition")):
        assert rel_maxnorm(x, ref) <= 1e-12  # reassociated sums (band 200: the partition is a candidate)
    else:
        assert np.array_equal(x.view(np.uint64), ref.view(np.uint64)), rel_maxnorm(x, ref)
    plan.close()


@pytest.mark.parametrize("kind", ["upper", "unit", "upper_unit"])
def test_triangular_kinds_partitioned(cuda, kind):
    """The partitioned solve on the converted matrix (bandwidth 100, 4000
    rows: 31 blocks), one right-hand side through solve and four through
    reset / launch."""
    import torch
    rp, ci, v = random_lower(4000, max_off=6, band=100, seed=9, dup=True)
    upper = kind in ("upper", "upper_unit")
    unit = kind in ("unit", "upper_unit")
    if upper:
        rp, ci, v = upper_from_lower(rp, ci, v)
    if unit:
        rp, ci, v = drop_diagonal(rp, ci, v, upper=upper)
    n = len(rp) - 1
    plan = _plan(cuda, rp, ci, v, algo="partition", upper=upper, unit_diagonal=unit)
    assert plan.info()["algo"] == "partition"
    b = api.uniform(45, n)
    x = plan.solve(torch.from_numpy(b).to(cuda)).cpu().numpy()
    assert plan.status() == 0
    assert rel_maxnorm(x, oracle.csr_solve(rp, ci, v, b, upper=upper, unit=unit, fused=True)) <= 1e-12
    k = 4
    b4 = api.uniform(46, n * k)
    bt = torch.from_numpy(b4.reshape(n, k)).to(cuda)
    x4 = torch.empty_like(bt)
    plan.reset(x4)
    plan.launch(bt, x4)
    torch.cuda.synchronize()
    assert plan.status() == 0
    ref4 = oracle.csr_solve(rp, ci, v, b4, rhs=k, upper=upper, unit=unit, fused=True)
    assert rel_maxnorm(x4.cpu().numpy().reshape(-1), ref4) <= 1e-12
    plan.close()


@pytest.mark.parametrize("algo", ["wrtfst", "local"])
def test_upper_multi_rhs_and_reset_launch(cuda, algo):
    import torch
    rp, ci, v, upper, unit = _mat("upper", n=2000, seed=5)
    n, k = len(rp) - 1, 4
    b = api.uniform(42, n * k)
    plan = _plan(cuda, rp, ci, v, algo=algo, upper=True)
    bt = torch.from_numpy(b.reshape(n, k)).to(cuda)
    x = torch.empty_like(bt)
    plan.reset(x)
    plan.launch(bt, x)
    torch.cuda.synchronize()
    assert plan.status() == 0
    ref = oracle.csr_solve(rp, ci, v, b, rhs=k, upper=True, fused=True)
    assert np.array_equal(x.cpu().numpy().reshape(-1).view(np.uint64), ref.view(np.uint64))
    plan.close()


def test_one_shot_host_solves(cuda):
    for kind in ("upper", "unit", "upper_unit"):
        rp, ci, v, upper, unit = _mat(kind, n=1500, seed=6)
        b = api.uniform(43, len(rp) - 1)
        x = api.sptrsv_csr_solve(rp, ci, v, b, upper=upper, unit_diagonal=unit)
        ref = oracle.csr_solve(rp, ci, v, b, upper=upper, unit=unit, fused=True)
        assert np.array_equal(x.view(np.uint64), ref.view(np.uint64)), kind
    # FP32
    rp, ci, v, upper, unit = _mat("upper", n=1500, seed=7)
    v32, b32 = v.astype(np.float32), api.uniform(44, len(rp) - 1).astype(np.float32)
    x = api.sptrsv_csr_solve(rp, ci, v32, b32, upper=True)
    assert np.array_equal(x.view(np.uint32), oracle.csr_solve(rp, ci, v32, b32, upper=True).view(np.uint32))


def test_triangular_layout_errors(cuda):
    from capellinisptrsv_amd import CapError
    rp, ci, v, _, _ = _mat("upper", n=300, seed=8)
    # an upper row whose first entry is not its diagonal
    bad = ci.copy()
    r = 10
    bad[rp[r]], bad[rp[r] + 1] = bad[rp[r] + 1], bad[rp[r]]
    with pytest.raises(CapError) as e:
        _plan(cuda, rp, bad, v, upper=True)
    assert e.value.code == -5
    # a lower matrix declared upper: entries below the diagonal
    lrp, lci, lv = random_lower(300, max_off=3, band=20, seed=9)
    with pytest.raises(CapError) as e:
        _plan(cuda, lrp, lci, lv, upper=True)
    assert e.value.code in (-4, -5)
    # unit diagonal declared while a diagonal is stored
    with pytest.raises(CapError) as e:
        _plan(cuda, lrp, lci, lv, unit_diagonal=True)
    assert e.value.code == -4


def test_converted_plans_have_no_adjoint_or_refinement(cuda):
    import torch
    from capellinisptrsv_amd import CapError
    rp, ci, v, _, _ = _mat("upper", n=500, seed=10)
    plan = _plan(cuda, rp, ci, v, algo="wrtfst", upper=True)
    g = torch.ones(len(rp) - 1, dtype=torch.float64, device=cuda)
    with pytest.raises(CapError) as e:
        plan.adjoint(g, g)
    assert e.value.code == -10
    with pytest.raises(CapError) as e:
        plan.solve_refined(g)
    assert e.value.code == -10
