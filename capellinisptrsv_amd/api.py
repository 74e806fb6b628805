"""Host-side mirror of the reference's interface on numpy arrays.

Same names, argument meaning and return conventions as the reference's C
helpers and solve wrappers (an int return code where the reference returns
one, outputs in place of its pointer arguments), backed by
libcapellini_sptrsv.so:

  read_mtx                  CUDA/ourWrtFst/read_mtx.h:6-191
  change2tran               CUDA/ourWrtFst/read_mtx.h:194-257
  get_x_b                   CUDA/ourWrtFst/read_mtx.h:260-278
  matrix_layer              CUDA/ourWrtFst/tranpose.h:78-142
  matrix_transposition      CUDA/ourWrtFst/tranpose.h:36-75
  matrix_warp4              CUDA/mix/tranpose.h:1423-1477
  spts_syncfree_hip_csr_wrt CUDA/ourWrtFst/spts_syncfree_cuda_csr_wrt.h:66-242
  spts_syncfree_hip_csr     CUDA/our2Part/spts_syncfree_cuda_csr.h:133-299
  spts_syncfree_hip_mix     CUDA/mix/spts_syncfree_cuda_csr_wrt.h:146-352
  sptrsv_syncfree_hip_csc   CUDA/SyncFree_csc/spts_syncfree_cuda2.h:89-292

Precision follows the dtype of the value array (float64 -> the _d entry,
float32 -> the _s entry), the runtime counterpart of the VALUE_TYPE switch.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_double, c_int, POINTER

import numpy as np

from . import _native as N

__all__ = [
    "read_mtx", "change2tran", "get_x_b", "matrix_layer", "matrix_transposition", "matrix_warp4",
    "spts_syncfree_hip_csr_wrt", "spts_syncfree_hip_csr", "spts_syncfree_hip_mix", "sptrsv_syncfree_hip_csc",
    "sptrsv_csr_solve", "validate_csr", "generate", "uniform", "datasize", "make_config", "CONFIGS",
    "make_batch",
]


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _pi(a):
    return a.ctypes.data_as(POINTER(c_int))


def _pv(a):
    return a.ctypes.data_as(POINTER(c_double) if a.dtype == np.float64 else POINTER(ctypes.c_float))


def _suffix(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return "d"
    if dt == np.float32:
        return "s"
    raise TypeError(f"values must be float64 or float32, got {dt}")


def _take(ptr, count, dtype):
    """Copy a malloc'ed C array into numpy and free it."""
    L = N.lib()
    if count == 0:
        out = np.zeros(0, dtype=dtype)
    else:
        out = np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)
    L.cap_free(ctypes.cast(ptr, ctypes.c_void_p))
    return out


# --------------------------------------------------------------- host helpers
def read_mtx(filename: str, dtype=np.float64):
    """Returns (ret, m, n, nnzA, row_ptr, col_idx, val); arrays are None when ret != 0."""
    L = N.lib()
    sfx = _suffix(dtype)
    m, n, nnz = c_int(), c_int(), c_int()
    rp, ci = POINTER(c_int)(), POINTER(c_int)()
    v = POINTER(c_double)() if sfx == "d" else POINTER(ctypes.c_float)()
    ret = getattr(L, f"cap_read_mtx_{sfx}")(str(filename).encode(), byref(m), byref(n), byref(nnz),
                                             byref(rp), byref(ci), byref(v))
    if ret != 0:
        return ret, 0, 0, 0, None, None, None
    return (ret, m.value, n.value, nnz.value, _take(rp, m.value + 1, np.int32), _take(ci, nnz.value, np.int32),
            _take(v, nnz.value, np.dtype(dtype)))


def change2tran(m: int, nnzA: int, row_ptr, col_idx, val, keep_values: bool = False):
    """Returns (nnzL, row_ptr_L, col_idx_L, val_L)."""
    L = N.lib()
    val = np.ascontiguousarray(val)
    sfx = _suffix(val.dtype)
    rp, ci = _i32(row_ptr), _i32(col_idx)
    nnzL = c_int()
    orp, oci = POINTER(c_int)(), POINTER(c_int)()
    ov = POINTER(c_double)() if sfx == "d" else POINTER(ctypes.c_float)()
    N.check(getattr(L, f"cap_change2tran_{sfx}")(m, nnzA, _pi(rp), _pi(ci), _pv(val), int(keep_values),
                                                 byref(nnzL), byref(orp), byref(oci), byref(ov)), "change2tran")
    k = nnzL.value
    return k, _take(orp, m + 1, np.int32), _take(oci, k, np.int32), _take(ov, k, val.dtype)


def get_x_b(m: int, n: int, row_ptr, col_idx, val):
    """Returns (x_ref, b): x_ref = 1, b = A x_ref."""
    L = N.lib()
    val = np.ascontiguousarray(val)
    sfx = _suffix(val.dtype)
    rp, ci = _i32(row_ptr), _i32(col_idx)
    x = np.empty(n, dtype=val.dtype)
    b = np.empty(m, dtype=val.dtype)
    getattr(L, f"cap_get_x_b_{sfx}")(m, n, _pi(rp), _pi(ci), _pv(val), _pv(x), _pv(b))
    return x, b


def matrix_layer(m: int, n: int, nnz: int, row_ptr, col_idx, val=None, return_levels: bool = False):
    """Returns (layer, parallelism) [, per-row level]."""
    rp, ci = _i32(row_ptr), _i32(col_idx)
    lv = np.zeros(max(m, 1), dtype=np.int32)
    layer, par = c_int(), c_double()
    N.lib().cap_matrix_layer(m, _pi(rp), _pi(ci), byref(layer), byref(par), _pi(lv))
    if return_levels:
        return layer.value, par.value, lv[:m]
    return layer.value, par.value


def matrix_transposition(m: int, n: int, nnz: int, row_ptr, col_idx, val):
    """Returns (csc_row_idx, csc_col_ptr, csc_val)."""
    val = np.ascontiguousarray(val)
    sfx = _suffix(val.dtype)
    rp, ci = _i32(row_ptr), _i32(col_idx)
    cr = np.empty(max(nnz, 1), np.int32)
    cp = np.empty(n + 1, np.int32)
    cv = np.empty(max(nnz, 1), val.dtype)
    getattr(N.lib(), f"cap_transpose_{sfx}")(m, n, nnz, _pi(rp), _pi(ci), _pv(val), _pi(cr), _pi(cp), _pv(cv))
    return cr[:nnz], cp, cv[:nnz]


def matrix_warp4(m: int, n: int, nnz: int, row_ptr, col_idx=None, val=None, border: int = 16, chunk: int = 64):
    """Returns (Len, warp_num[:Len], warp_occupy, element_occupy)."""
    rp = _i32(row_ptr)
    wn = np.zeros(m + 1, np.int32)
    ln, wo, eo = c_int(), c_double(), c_double()
    N.lib().cap_matrix_warp4(m, nnz, _pi(rp), int(border), int(chunk), byref(ln), _pi(wn), byref(wo), byref(eo))
    return ln.value, wn[:ln.value].copy(), wo.value, eo.value


def validate_csr(n: int, nnz: int, row_ptr, col_idx, val):
    """Returns (code, bad_row) of the solve-layout check."""
    val = np.ascontiguousarray(val)
    sfx = _suffix(val.dtype)
    rp, ci = _i32(row_ptr), _i32(col_idx)
    bad = c_int(-1)
    code = getattr(N.lib(), f"cap_validate_csr_{sfx}")(n, nnz, _pi(rp), _pi(ci), _pv(val), byref(bad))
    return code, bad.value


# ------------------------------------------------------------- serial CPU solves
FORWARD, BACKWARD = 0, 1  # SUBSTITUTION_FORWARD / _BACKWARD (openCL/SynFree_csc/common.h:25-26)


def serial_analyser(csc_row_idx, m: int, nnz: int):
    """In-degree of every row of a CSC matrix (sptrsv_syncfree_serialref.h:6-23)."""
    deg = np.zeros(max(m, 1), np.int32)
    N.check(N.lib().cap_serial_analyser(_pi(_i32(csc_row_idx)), m, m, nnz, _pi(deg)), "cap_serial_analyser")
    return deg[:m]


def serial_executor(csc_col_ptr, csc_row_idx, csc_val, in_degree, m: int, n: int, substitution: int, rhs: int, b):
    """Column-driven serial substitution on the CPU (serialref.h:25-108):
    returns x (n*rhs values, interleaved); raises CapError on a malformed matrix."""
    cv = np.ascontiguousarray(csc_val)
    sfx = _suffix(cv.dtype)
    b = np.ascontiguousarray(b, dtype=cv.dtype).reshape(-1)
    if b.size != n * rhs:
        raise ValueError("b must hold n*rhs values")
    x = np.zeros(max(n * rhs, 1), cv.dtype)
    N.check(getattr(N.lib(), f"cap_serial_executor_{sfx}")(_pi(_i32(csc_col_ptr)), _pi(_i32(csc_row_idx)), _pv(cv),
                                                           _pi(_i32(in_degree)), m, n, substitution, rhs, _pv(b),
                                                           _pv(x)), "cap_serial_executor")
    return x[:n * rhs]


def sptrsv_syncfree_serialref(csc_col_ptr, csc_row_idx, csc_val, m: int, n: int, nnz: int, substitution: int,
                              rhs: int, b, x_ref):
    """The reference's timed serial wrapper (serialref.h:110-177): returns (ret, x)."""
    cv = np.ascontiguousarray(csc_val)
    sfx = _suffix(cv.dtype)
    b = np.ascontiguousarray(b, dtype=cv.dtype).reshape(-1)
    xr = np.ascontiguousarray(x_ref, dtype=cv.dtype).reshape(-1)
    if m == n and (b.size != n * rhs or xr.size != n * rhs):  # a non-square call returns -1 untouched
        raise ValueError("b and x_ref must hold n*rhs values")
    x = np.zeros(max(n * rhs, 1), cv.dtype)
    ret = getattr(N.lib(), f"sptrsv_syncfree_serialref_{sfx}")(_pi(_i32(csc_col_ptr)), _pi(_i32(csc_row_idx)),
                                                                _pv(cv), m, n, nnz, substitution, rhs, _pv(x),
                                                                _pv(b), _pv(xr))
    return ret, x[:n * rhs]


def serial_csr_solve(row_ptr, col_idx, val, b, rhs: int = 1, arith: str = "reference"):
    """Row-ordered serial forward substitution on the CPU over the solve
    contract's CSR, terms in stored order; arith "reference" (rounded product,
    rounded sum) or "fused" (the GPU kernels' default)."""
    val = np.ascontiguousarray(val)
    sfx = _suffix(val.dtype)
    n = len(row_ptr) - 1
    b = np.ascontiguousarray(b, dtype=val.dtype).reshape(-1)
    if b.size != n * rhs:
        raise ValueError("b must hold n*rhs values")
    x = np.zeros(max(n * rhs, 1), val.dtype)
    N.check(getattr(N.lib(), f"cap_serial_csr_{sfx}")(n, len(col_idx), _pi(_i32(row_ptr)), _pi(_i32(col_idx)),
                                                      _pv(val), rhs, _pv(b), _pv(x), N.ARITH[arith]),
            "cap_serial_csr")
    return x[:n * rhs]


def datasize(m: int, n: int, nnzL: int, value_bytes: int, len_warp_num: int = 0) -> float:
    return N.lib().cap_datasize(m, n, nnzL, value_bytes, len_warp_num)


# ---------------------------------------------------- benchmark-compatible solves
def spts_syncfree_hip_csr_wrt(row_ptr, col_idx, val, m: int, n: int, nnzL: int):
    """Returns (ret, solve_time_ms, gflops, bandwith)."""
    return _bench("spts_syncfree_hip_csr_wrt", row_ptr, col_idx, val, m, n, nnzL, mix=False)


def spts_syncfree_hip_csr(row_ptr, col_idx, val, m: int, n: int, nnzL: int):
    """our2Part.  Returns (ret, solve_time_ms, gflops, bandwith)."""
    return _bench("spts_syncfree_hip_csr", row_ptr, col_idx, val, m, n, nnzL, mix=False)


def spts_syncfree_hip_mix(row_ptr, col_idx, val, m: int, n: int, nnzL: int):
    """Returns (ret, solve_time_ms, gflops, bandwith, pre_time_ms, warp_occupy, element_occupy)."""
    return _bench("spts_syncfree_hip_mix", row_ptr, col_idx, val, m, n, nnzL, mix=True)


def spts_syncfree_hip_csr_algo(row_ptr, col_idx, val, m: int, n: int, nnzL: int, algo: str = "auto"):
    """The benchmark-compatible wrapper with a chosen algorithm (auto, wrtfst,
    twopart, csc, local, partition, level).  Returns (ret, solve_time_ms, gflops, bandwith)."""
    if algo not in N.ALGO or algo == "mix":
        raise ValueError(f"algorithm {algo!r}: one of auto, wrtfst, twopart, csc, local, partition, level")
    return _bench("spts_syncfree_hip_csr_algo", row_ptr, col_idx, val, m, n, nnzL, mix=False, algo=N.ALGO[algo])


def _bench(name, row_ptr, col_idx, val, m, n, nnzL, mix, algo=None):
    val = np.ascontiguousarray(val)
    sfx = _suffix(val.dtype)
    rp, ci = _i32(row_ptr), _i32(col_idx)
    out = [c_double() for _ in range(6 if mix else 3)]
    extra = [] if algo is None else [algo]
    ret = getattr(N.lib(), f"{name}_{sfx}")(_pi(rp), _pi(ci), _pv(val), m, n, nnzL, *extra, *[byref(o) for o in out])
    return (ret, *[o.value for o in out])


def sptrsv_syncfree_hip_csc(csc_col_ptr, csc_row_idx, csc_val, m: int, n: int, nnz: int, b, x_ref):
    """Returns (ret, x, pre_time_ms, solve_time_ms, gflops, bandwidth)."""
    cv = np.ascontiguousarray(csc_val)
    sfx = _suffix(cv.dtype)
    cp, ri = _i32(csc_col_ptr), _i32(csc_row_idx)
    b = np.ascontiguousarray(b, dtype=cv.dtype)
    xr = np.ascontiguousarray(x_ref, dtype=cv.dtype)
    x = np.zeros(n, dtype=cv.dtype)
    out = [c_double() for _ in range(4)]
    ret = getattr(N.lib(), f"sptrsv_syncfree_hip_csc_{sfx}")(_pi(cp), _pi(ri), _pv(cv), m, n, nnz, _pv(x), _pv(b),
                                                              _pv(xr), *[byref(o) for o in out])
    return (ret, x, *[o.value for o in out])


def sptrsv_csr_solve(row_ptr, col_idx, val, b, algo: str = "wrtfst", sync: str = "value", border: int = 0,
                     spin_limit: int = 0, upper: bool = False, unit_diagonal: bool = False, arith: str = "fused"):
    """One-shot host solve T x = b on the GPU; raises CapError on failure.  T is
    lower triangular with the diagonal last in each row, or with upper=True
    upper triangular with the diagonal first; unit_diagonal=True: no diagonal
    is stored and it is taken as 1.  arith="reference": each term rounded as
    the reference's serial executor rounds it (product, then sum)."""
    val = np.ascontiguousarray(val)
    sfx = _suffix(val.dtype)
    rp, ci = _i32(row_ptr), _i32(col_idx)
    n = len(rp) - 1
    b = np.ascontiguousarray(b, dtype=val.dtype)
    x = np.zeros(n, dtype=val.dtype)
    o = N.default_opts(algo=algo, sync=sync, border=border, spin_limit=spin_limit, upper=int(upper),
                       unit_diag=int(unit_diagonal), arith=arith)
    N.check(getattr(N.lib(), f"sptrsv_csr_solve_{sfx}")(n, int(rp[-1]), _pi(rp), _pi(ci), _pv(val), _pv(b), _pv(x),
                                                         byref(o)), "sptrsv_csr_solve")
    return x


# ----------------------------------------------------------------- generators
_KIND = {"band": 0, "level": 1, "bernoulli": 2, "band_rowcount": 3}


def generate(kind: str, n: int = 0, offdiag: int = 0, band: int = 0, levels: int = 0, width: int = 0,
             depth: int = 0, p: float = 0.0, seed: int = 0, dtype=np.float64):
    """Seeded synthetic lower-triangular CSR (see cap_generate_d).  Returns (row_ptr, col_idx, val)."""
    g = N.cap_gen_params(_KIND[kind], n, offdiag, band, levels, width, depth, p, seed)
    sfx = _suffix(dtype)
    L = N.lib()
    nnz = c_int()
    rp = np.zeros(n + 1, np.int32)
    N.check(getattr(L, f"cap_generate_{sfx}")(byref(g), byref(nnz), _pi(rp), None, None), "generate")
    ci = np.zeros(max(nnz.value, 1), np.int32)
    v = np.zeros(max(nnz.value, 1), np.dtype(dtype))
    N.check(getattr(L, f"cap_generate_{sfx}")(byref(g), byref(nnz), _pi(rp), _pi(ci), _pv(v)), "generate")
    return rp, ci[:nnz.value], v[:nnz.value]


def uniform(seed: int, count: int, lo: float = -1.0, hi: float = 1.0):
    out = np.empty(count, np.float64)
    N.lib().cap_uniform_d(seed, count, lo, hi, out.ctypes.data_as(POINTER(c_double)))
    return out


# The benchmark workloads (sizes as specified for the reference comparison).
CONFIGS = {
    0: dict(kind="band", n=100_000, offdiag=3, band=1000),
    1: dict(kind="level", n=2_000_000, levels=200, width=10_000, depth=3, offdiag=2),
    2: dict(kind="bernoulli", n=500_000, band=128, p=0.3),
    4: dict(kind="level", n=1_000_000, levels=100, width=10_000, depth=3, offdiag=2),
}


def make_config(cfg: int, seed: int = 0, dtype=np.float64, scale: float = 1.0):
    """Matrix of workload `cfg` (0, 1, 2, 4), optionally shrunk by `scale` for tests."""
    c = dict(CONFIGS[cfg])
    if scale != 1.0:
        if c["kind"] == "level":
            c["levels"] = max(2, int(c["levels"] * scale))
            c["n"] = c["levels"] * c["width"]
        else:
            c["n"] = max(2, int(c["n"] * scale))
    return generate(seed=seed, dtype=dtype, **c)


def make_batch(n_mat: int = 256, seed: int = 0, n_lo: int = 2000, n_hi: int = 200_000, band_lo: int = 64,
               band_hi: int = 4096, max_row_nnz: int = 16, dtype=np.float64):
    """The batched workload: n_mat matrices with n log-uniform in [n_lo, n_hi],
    nnz/row (diagonal included) uniform in {1..max_row_nnz} and the band
    uniform in [band_lo, band_hi], each by the band rule of workload 0.

    Returns (mats, row_off, row_ptr, col_idx, values): the list of per-matrix
    CSR triples and their concatenation in the batch layout (global row_ptr,
    column indices local to each matrix)."""
    rng = np.random.default_rng(seed)
    mats = []
    for m in range(n_mat):
        nm = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        band = int(rng.integers(band_lo, band_hi + 1))
        mats.append(generate("band_rowcount", n=nm, offdiag=max_row_nnz, band=band, seed=1000 + m, dtype=dtype))
    row_off = np.cumsum([0] + [len(m[0]) - 1 for m in mats]).astype(np.int32)
    nnz_off = np.cumsum([0] + [len(m[1]) for m in mats])
    rp = np.concatenate([m[0][:-1] + o for m, o in zip(mats, nnz_off)] + [[nnz_off[-1]]]).astype(np.int32)
    ci = np.concatenate([m[1] for m in mats]).astype(np.int32)
    v = np.concatenate([m[2] for m in mats])
    return mats, row_off, rp, ci, v
