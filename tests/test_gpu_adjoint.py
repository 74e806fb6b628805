"""GPU adjoint (grad_b = L^{-T} g, grad_val = -grad_b[row] * x[col]) through the C ABI.

Checks, FP64:
  * against the oracle's column-driven backward executor on the CSR of L
    (relative max-norm 1e-12: the kernel's atomic left sums arrive in a
    scheduling-dependent order);
  * against a dense FP64 torch autograd gradient of 0.5*||L^{-1}b - x_t||^2 on
    an n = 2,000 instance of the level generator (width 20, 100 levels);
  * against central finite differences (h = 1e-6) on 64 random value entries
    and 16 entries of b of the n = 1,000,000 level matrix.
"""
import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import api
from conftest import random_lower, rel_maxnorm

pytestmark = pytest.mark.gpu


def _plan(cuda, rp, ci, v, algo="wrtfst", adjoint="auto"):
    import torch
    from capellinisptrsv_amd import Plan
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v)]
    return Plan(*d, algo=algo, sync="value", adjoint=adjoint), d


@pytest.mark.parametrize("mode", ["gather", "scatter"])
@pytest.mark.parametrize("mat", ["band", "level", "chain_dup"])
def test_adjoint_matches_oracle(cuda, mat, mode):
    import torch
    if mat == "band":
        rp, ci, v = random_lower(4000, max_off=6, band=60, seed=21)
    elif mat == "level":
        rp, ci, v = api.generate("level", n=50_000, levels=50, width=1000, depth=3, offdiag=2, seed=3)
    else:
        rp, ci, v = random_lower(3000, max_off=5, band=10, seed=22, dup=True)
    n = len(rp) - 1
    plan, d = _plan(cuda, rp, ci, v, adjoint=mode)
    b = api.uniform(1, n)
    g = api.uniform(2, n)
    x = plan.solve(torch.from_numpy(b).to(cuda))
    gb, gv = plan.adjoint(torch.from_numpy(g).to(cuda), x)
    torch.cuda.synchronize()
    assert plan.status() == 0
    y_ref = oracle.adjoint(rp, ci, v, g, fused=False)
    assert rel_maxnorm(gb.cpu().numpy(), y_ref) <= 1e-12
    if mode == "gather":  # the executor's backward order, fused: bit-exact
        assert np.array_equal(gb.cpu().numpy(), oracle.adjoint(rp, ci, v, g, fused=True))
    gv_ref = oracle.grad_val(rp, ci, y_ref, oracle.csr_forward(rp, ci, v, b, fused=True))
    assert rel_maxnorm(gv.cpu().numpy(), gv_ref) <= 1e-12
    # the oracle's adjoint is the executor's BACKWARD mode on L read as CSC of L^T
    cp, ri, cv = rp, ci, v
    indeg = np.bincount(ci, minlength=n).astype(np.int32)
    assert np.array_equal(oracle.csc_executor(cp, ri, cv, indeg, n, n, oracle.BACKWARD, 1, g), y_ref)


def test_adjoint_against_dense_autograd_n2000(cuda):
    import torch
    from capellinisptrsv_amd import SpTRSV
    from capellinisptrsv_amd import Plan
    rp, ci, v = api.generate("level", n=2000, levels=100, width=20, depth=3, offdiag=2, seed=7)
    n = len(rp) - 1
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v)]
    vals = d[2].clone().requires_grad_(True)
    # the plan reads the parameter's storage in place
    plan = Plan(d[0], d[1], vals.detach(), algo="wrtfst", sync="value")
    b = torch.from_numpy(api.uniform(3, n)).to(cuda).requires_grad_(True)
    xt = torch.from_numpy(api.uniform(4, n)).to(cuda)
    x = SpTRSV.apply(vals, b, plan)
    loss = 0.5 * ((x - xt) ** 2).sum()
    loss.backward()
    # dense FP64 reference on the CPU
    rows = np.repeat(np.arange(n), np.diff(rp))
    Ld = torch.zeros(n, n, dtype=torch.float64)
    vd = torch.from_numpy(v.copy()).requires_grad_(True)
    Ld = Ld.index_put((torch.from_numpy(rows), torch.from_numpy(ci.astype(np.int64))), vd, accumulate=True)
    bd = torch.from_numpy(api.uniform(3, n)).requires_grad_(True)
    xd = torch.linalg.solve_triangular(Ld, bd.unsqueeze(1), upper=False).squeeze(1)
    lossd = 0.5 * ((xd - xt.cpu()) ** 2).sum()
    lossd.backward()
    assert rel_maxnorm(x.detach().cpu().numpy(), xd.detach().numpy()) <= 1e-12
    assert rel_maxnorm(b.grad.cpu().numpy(), bd.grad.numpy()) <= 1e-10
    assert rel_maxnorm(vals.grad.cpu().numpy(), vd.grad.numpy()) <= 1e-10


def test_adjoint_finite_differences_n1m(cuda):
    import torch
    from capellinisptrsv_amd import Plan
    rp, ci, v = api.make_config(4, seed=0)
    n = len(rp) - 1
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v)]
    plan = Plan(*d, algo="wrtfst", sync="value")
    b = torch.from_numpy(api.uniform(5, n)).to(cuda)
    xt = torch.from_numpy(api.uniform(6, n)).to(cuda)
    x = plan.solve(b)
    gb, gv = plan.adjoint((x - xt).contiguous(), x)
    torch.cuda.synchronize()
    h = 1e-6
    rng = np.random.default_rng(0)

    def dloss(xp, xm):  # loss(+) - loss(-) without cancellation
        return float((0.5 * (xp - xm) * (xp + xm - 2 * xt)).sum())

    for k in rng.choice(len(ci), size=64, replace=False):
        vp, vm = d[2].clone(), d[2].clone()
        vp[k] += h
        vm[k] -= h
        pp = Plan(d[0], d[1], vp, algo="wrtfst", sync="value", validate=False)
        pm = Plan(d[0], d[1], vm, algo="wrtfst", sync="value", validate=False)
        fd = dloss(pp.solve(b), pm.solve(b)) / (2 * h)
        an = float(gv[k])
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (k, fd, an)
        pp.close()
        pm.close()
    for i in rng.choice(n, size=16, replace=False):
        bp, bm = b.clone(), b.clone()
        bp[i] += h
        bm[i] -= h
        fd = dloss(plan.solve(bp), plan.solve(bm)) / (2 * h)
        an = float(gb[i])
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (i, fd, an)
