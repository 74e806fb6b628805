"""The allocation-free solve (sptrsv_csr_solve_noalloc_d) through the C ABI.

It runs the one-shot solve's write-first kernel on the caller's device arrays
with its tickets and watchdog word in the library's static per-device scratch
array.  Bar: bit-identical to the existing one-shot entry point
(sptrsv_csr_solve_d, host arrays, same kernel) and to the oracle's fused
sequential forward substitution, on two calls in a row -- the second with a
smaller matrix, so the scratch's leftover tickets from the first call must be
cleared -- and again after an unrelated plan solve on the same device.
"""
import ctypes

import numpy as np
import pytest

import oracle
from capellinisptrsv_amd import _native as N
from capellinisptrsv_amd import api
from conftest import random_lower

pytestmark = pytest.mark.gpu


def _existing(rp, ci, v, b):
    """The existing one-shot entry point (host arrays), same kernel."""
    x = np.zeros(len(rp) - 1)
    o = N.default_opts(algo="wrtfst", sync="value")
    P = ctypes.POINTER
    rc = N.lib().sptrsv_csr_solve_d(len(rp) - 1, len(ci), rp.ctypes.data_as(P(ctypes.c_int)),
                                    ci.ctypes.data_as(P(ctypes.c_int)), v.ctypes.data_as(P(ctypes.c_double)),
                                    b.ctypes.data_as(P(ctypes.c_double)), x.ctypes.data_as(P(ctypes.c_double)),
                                    ctypes.byref(o))
    assert rc == 0
    return x


def _noalloc(cuda, rp, ci, v, b):
    import torch
    d = [torch.from_numpy(a).to(cuda) for a in (rp, ci, v, b)]
    x = torch.empty(len(rp) - 1, dtype=torch.float64, device=cuda)
    stream = torch.cuda.current_stream()
    rc = N.lib().sptrsv_csr_solve_noalloc_d(len(rp) - 1, len(ci), *(ctypes.c_void_p(t.data_ptr()) for t in d),
                                            ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(stream.cuda_stream))
    assert rc == 0
    stream.synchronize()
    return x.cpu().numpy()


def test_noalloc_matches_the_one_shot_solve_on_two_calls(cuda):
    cases = [api.make_config(1, seed=4, scale=0.05),          # 100,000 rows, wide levels
             random_lower(3000, max_off=5, band=40, seed=8)]  # then a smaller band
    for rp, ci, v in cases:
        b = api.uniform(17, len(rp) - 1)
        got = _noalloc(cuda, rp, ci, v, b)
        assert np.array_equal(got.view(np.uint64), _existing(rp, ci, v, b).view(np.uint64))
        assert np.array_equal(got, oracle.csr_forward(rp, ci, v, b, fused=True))


def test_noalloc_after_a_plan_solve_and_argument_checks(cuda):
    import torch
    from capellinisptrsv_amd import Plan
    rp, ci, v = api.make_config(0, seed=6, scale=0.2)
    b = api.uniform(18, len(rp) - 1)
    plan = Plan(*(torch.from_numpy(a).to(cuda) for a in (rp, ci, v)), algo="mix")
    plan.solve(torch.from_numpy(b).to(cuda))
    plan.close()
    got = _noalloc(cuda, rp, ci, v, b)
    assert np.array_equal(got, oracle.csr_forward(rp, ci, v, b, fused=True))
    L = N.lib()
    null = ctypes.c_void_p(0)
    assert L.sptrsv_csr_solve_noalloc_d(-1, 0, null, null, null, null, null, null) == N.CAP_ERR_ARG
    assert L.sptrsv_csr_solve_noalloc_d(10, 5, null, null, null, null, null, null) == N.CAP_ERR_ARG
    assert L.sptrsv_csr_solve_noalloc_d(0, 0, ctypes.c_void_p(1), null, null, null, null, null) == 0
