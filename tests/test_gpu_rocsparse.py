"""The rocSPARSE `csrsv` comparator against the oracle.

The reference validates its vendor baseline inline (CUDA/cuSP/
spts_cuSPARSE.h:191-204: x against x_ref) for the level-free and the
level-scheduled cuSPARSE solves.  Here rocSPARSE's csrsv (analysis once,
then solves) runs on the same matrices as this library's kernels and must
agree with the oracle's sequential forward substitution: FP64 within 1e-12
relative max-norm, FP32 within 1e-5 of the FP64 oracle.  Repeated solves
after one analysis give the same bits.
"""
import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import api
from conftest import random_lower, rel_maxnorm

pytestmark = pytest.mark.gpu


def _cases():
    yield "band", random_lower(3000, max_off=5, band=40, seed=31)
    yield "cfg0_small", api.make_config(0, seed=2, scale=0.2)
    yield "cfg1_small", api.make_config(1, seed=2, scale=0.02)
    yield "cfg2_small", api.make_config(2, seed=2, scale=0.01)


CASES = list(_cases())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name,mat", CASES, ids=[c[0] for c in CASES])
def test_rocsparse_csrsv_matches_oracle(cuda, name, mat, dtype):
    import torch
    from capellinisptrsv_amd.rocsparse_cmp import Csrsv
    rp, ci, v = mat
    n = len(rp) - 1
    b = api.uniform(32, n)
    ref = oracle.csr_forward(rp, ci, v, b, fused=True)
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v.astype(dtype))]
    rs = Csrsv(*d)
    rs.analysis()
    bt = torch.from_numpy(b.astype(dtype)).to(cuda)
    x1 = rs.solve(bt, torch.empty_like(bt)).cpu().numpy()
    x2 = rs.solve(bt, torch.empty_like(bt)).cpu().numpy()
    rs.close()
    assert np.array_equal(x1, x2), name
    tol = 1e-12 if dtype == np.float64 else 1e-5
    assert rel_maxnorm(x1, ref) <= tol, (name, rel_maxnorm(x1, ref))
