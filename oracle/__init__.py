"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and of the
reference-helper harness (oracle/_ref/libref_*.so) -- TEST INFRASTRUCTURE
ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's CPU baseline
leg, always as the checker, never as the thing measured or shipped.

Every oracle function restates the reference lines cited in oracle/oracle.h.
"""
from __future__ import annotations

import ctypes
import os
import time
from ctypes import POINTER, byref, c_double, c_float, c_int, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

PI, PD, PF = POINTER(c_int), POINTER(c_double), POINTER(c_float)
FORWARD, BACKWARD = 0, 1

_orc = None
_refs: dict = {}


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _pi(a):
    return a.ctypes.data_as(PI)


def _pv(a):
    return a.ctypes.data_as(PD if a.dtype == np.float64 else PF)


def _sfx(dtype):
    return "d" if np.dtype(dtype) == np.float64 else "s"


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is not built (make -C oracle oracle)")
        _orc = ctypes.CDLL(ORACLE_SO)
        _orc.orc_datasize.restype = c_double
        _orc.orc_gflops.restype = c_double
        _orc.orc_matrix_layer.restype = c_int
    return _orc


def _take(L, ptr, count, dtype):
    out = np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True) if count else np.zeros(0, dtype)
    L.orc_free(ctypes.cast(ptr, c_void_p))
    return out


# ------------------------------------------------------------------ oracle
def read_mtx(path, dtype=np.float64):
    L = lib()
    s = _sfx(dtype)
    m, n, nnz = c_int(), c_int(), c_int()
    rp, ci = PI(), PI()
    v = PD() if s == "d" else PF()
    ret = getattr(L, f"orc_read_mtx_{s}")(str(path).encode(), byref(m), byref(n), byref(nnz), byref(rp), byref(ci),
                                          byref(v))
    if ret != 0:
        return ret, 0, 0, 0, None, None, None
    return (0, m.value, n.value, nnz.value, _take(L, rp, m.value + 1, np.int32), _take(L, ci, nnz.value, np.int32),
            _take(L, v, nnz.value, np.dtype(dtype)))


def change2tran(m, nnzA, rp, ci, v):
    """Note: sorts rp/ci/v's rows in place like the reference (pass copies)."""
    L = lib()
    v = np.ascontiguousarray(v)
    s = _sfx(v.dtype)
    rp, ci = _i32(rp).copy(), _i32(ci).copy()
    v = v.copy()
    nnzL = c_int()
    orp, oci = PI(), PI()
    ov = PD() if s == "d" else PF()
    getattr(L, f"orc_change2tran_{s}")(m, nnzA, _pi(rp), _pi(ci), _pv(v), byref(nnzL), byref(orp), byref(oci),
                                       byref(ov))
    k = nnzL.value
    return k, _take(L, orp, m + 1, np.int32), _take(L, oci, k, np.int32), _take(L, ov, k, v.dtype)


def get_x_b(m, n, rp, ci, v):
    v = np.ascontiguousarray(v)
    x = np.empty(n, v.dtype)
    b = np.empty(m, v.dtype)
    getattr(lib(), f"orc_get_x_b_{_sfx(v.dtype)}")(m, n, _pi(_i32(rp)), _pi(_i32(ci)), _pv(v), _pv(x), _pv(b))
    return x, b


def matrix_layer(m, rp, ci):
    lv, par = c_int(), c_double()
    lib().orc_matrix_layer(m, _pi(_i32(rp)), _pi(_i32(ci)), byref(lv), byref(par))
    return lv.value, par.value


def row_levels(m, rp, ci):
    out = np.zeros(max(m, 1), np.int32)
    lib().orc_row_levels(m, _pi(_i32(rp)), _pi(_i32(ci)), _pi(out))
    return out[:m]


def transpose(m, n, nnz, rp, ci, v):
    v = np.ascontiguousarray(v)
    cr = np.zeros(max(nnz, 1), np.int32)
    cp = np.zeros(n + 1, np.int32)
    cv = np.zeros(max(nnz, 1), v.dtype)
    getattr(lib(), f"orc_transpose_{_sfx(v.dtype)}")(m, n, nnz, _pi(_i32(rp)), _pi(_i32(ci)), _pv(v), _pi(cr),
                                                      _pi(cp), _pv(cv))
    return cr[:nnz], cp, cv[:nnz]


def matrix_warp4(m, nnz, rp, border, chunk):
    wn = np.zeros(m + 1, np.int32)
    ln, wo, eo = c_int(), c_double(), c_double()
    lib().orc_matrix_warp4(m, nnz, _pi(_i32(rp)), border, chunk, byref(ln), _pi(wn), byref(wo), byref(eo))
    return ln.value, wn[:ln.value].copy(), wo.value, eo.value


def csc_analyser(csc_row, m, nnz):
    deg = np.zeros(max(m, 1), np.int32)
    lib().orc_csc_analyser(_pi(_i32(csc_row)), m, nnz, _pi(deg))
    return deg[:m]


def csc_executor(cp, ri, cv, indeg, m, n, substitution, rhs, b):
    cv = np.ascontiguousarray(cv)
    b = np.ascontiguousarray(b, dtype=cv.dtype)
    x = np.zeros(n * rhs, cv.dtype)
    getattr(lib(), f"orc_csc_executor_{_sfx(cv.dtype)}")(_pi(_i32(cp)), _pi(_i32(ri)), _pv(cv), _pi(_i32(indeg)), m,
                                                          n, substitution, rhs, _pv(b), _pv(x))
    return x


def csr_solve(rp, ci, v, b, rhs=1, upper=False, unit=False, fused=True):
    """Lower/upper, stored/implicit-unit-diagonal CSR solve in stored order (oracle.c)."""
    v = np.ascontiguousarray(v)
    b = np.ascontiguousarray(b, dtype=v.dtype).reshape(-1)
    n = len(rp) - 1
    x = np.zeros(n * rhs, dtype=v.dtype)
    getattr(lib(), f"orc_csr_solve_{_sfx(v.dtype)}")(n, _pi(_i32(rp)), _pi(_i32(ci)), _pv(v), rhs, _pv(b), _pv(x),
                                                     int(upper), int(unit), int(fused))
    return x


def csr_forward(rp, ci, v, b, rhs=1, fused=True):
    v = np.ascontiguousarray(v)
    n = len(rp) - 1
    b = np.ascontiguousarray(b, dtype=v.dtype).reshape(-1)
    x = np.zeros(n * rhs, v.dtype)
    getattr(lib(), f"orc_csr_forward_{_sfx(v.dtype)}")(n, _pi(_i32(rp)), _pi(_i32(ci)), _pv(v), rhs, _pv(b), _pv(x),
                                                        int(fused))
    return x


def adjoint(rp, ci, v, g, fused=True):
    n = len(rp) - 1
    v = np.ascontiguousarray(v, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    y = np.zeros(n, np.float64)
    lib().orc_adjoint_d(n, _pi(_i32(rp)), _pi(_i32(ci)), _pv(v), _pv(g), _pv(y), int(fused))
    return y


def grad_val(rp, ci, y, x):
    n = len(rp) - 1
    gv = np.zeros(len(ci), np.float64)
    lib().orc_grad_val_d(n, _pi(_i32(rp)), _pi(_i32(ci)), _pv(np.ascontiguousarray(y, np.float64)),
                         _pv(np.ascontiguousarray(x, np.float64)), _pv(gv))
    return gv


def residual(rp, ci, v, b, x, rhs=1):
    n = len(rp) - 1
    r = np.zeros(n * rhs, np.float64)
    lib().orc_residual_d(n, _pi(_i32(rp)), _pi(_i32(ci)), _pv(np.ascontiguousarray(v, np.float64)), rhs,
                         _pv(np.ascontiguousarray(b, np.float64).reshape(-1)),
                         _pv(np.ascontiguousarray(x, np.float64).reshape(-1)), _pv(r))
    return r


def refine2(rp, ci, v, b, rhs=1, rcp=True):
    """Two refinement steps (FP32 solve, FP64 residual, FP32 correction):
    returns (x0, d, x) with x = x0 + d in FP64.  rcp: the FP32 rows end with a
    product by the rounded reciprocal diagonal (the fused sweep) instead of a
    division (the one-workgroup plans' two separate solves)."""
    n = len(rp) - 1
    x0 = np.zeros(n * rhs, np.float32)
    d = np.zeros(n * rhs, np.float32)
    x = np.zeros(n * rhs, np.float64)
    lib().orc_refine2_d(n, _pi(_i32(rp)), _pi(_i32(ci)), _pv(np.ascontiguousarray(v, np.float64)), rhs,
                        _pv(np.ascontiguousarray(b, np.float64).reshape(-1)), _pv(x0), _pv(d), _pv(x), int(rcp))
    return x0, d, x


def datasize(m, n, nnzL, vb, ln=0):
    return lib().orc_datasize(m, n, nnzL, vb, ln)


def time_forward(rp, ci, v, b, repeats=3, budget_s=20.0, fused=False):
    """Median wall time (s) of the serial forward substitution on one core."""
    times = []
    csr_forward(rp, ci, v, b, fused=fused)  # warm-up
    t_start = time.perf_counter()
    for _ in range(repeats):
        t0 = time.perf_counter()
        csr_forward(rp, ci, v, b, fused=fused)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s:
            break
    return float(np.median(times))


# --------------------------------------------- reference helpers (_ref)
def ref_available(variant="d64") -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"libref_{variant}.so"))


def ref(variant="d64"):
    """The reference's own host helpers compiled by oracle/Makefile (double/float x warp 64/32)."""
    if variant not in _refs:
        path = os.path.join(REF_DIR, f"libref_{variant}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle ref)")
        h = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        h.ref_matrix_layer.restype = c_int
        _refs[variant] = h
    return _refs[variant]


def ref_read_mtx(path, variant="d64"):
    R = ref(variant)
    dt = np.float64 if R.ref_value_bytes() == 8 else np.float32
    m, n, nnz = c_int(), c_int(), c_int()
    rp, ci = PI(), PI()
    v = PD() if dt == np.float64 else PF()
    ret = R.ref_read_mtx(str(path).encode(), byref(m), byref(n), byref(nnz), byref(rp), byref(ci), byref(v))
    if ret != 0:
        return ret, 0, 0, 0, None, None, None

    def take(p, c, d):
        out = np.ctypeslib.as_array(p, shape=(c,)).astype(d, copy=True) if c else np.zeros(0, d)
        R.ref_free(ctypes.cast(p, c_void_p))
        return out
    return 0, m.value, n.value, nnz.value, take(rp, m.value + 1, np.int32), take(ci, nnz.value, np.int32), take(
        v, nnz.value, dt)


def ref_change2tran(m, nnzA, rp, ci, v, variant="d64"):
    """The reference allocates nnzA entries for L (read_mtx.h:198-199): callers
    must pass an A whose every row holds an entry with col >= row."""
    R = ref(variant)
    v = np.ascontiguousarray(v).copy()
    rp, ci = _i32(rp).copy(), _i32(ci).copy()
    nnzL = c_int()
    orp, oci = PI(), PI()
    ov = PD() if v.dtype == np.float64 else PF()
    R.ref_change2tran(m, nnzA, _pi(rp), _pi(ci), _pv(v), byref(nnzL), byref(orp), byref(oci), byref(ov))
    k = nnzL.value

    def take(p, c, d):
        out = np.ctypeslib.as_array(p, shape=(c,)).astype(d, copy=True) if c else np.zeros(0, d)
        R.ref_free(ctypes.cast(p, c_void_p))
        return out
    return k, take(orp, m + 1, np.int32), take(oci, k, np.int32), take(ov, k, v.dtype)


def ref_get_x_b(m, n, rp, ci, v, variant="d64"):
    v = np.ascontiguousarray(v)
    x = np.empty(n, v.dtype)
    b = np.empty(m, v.dtype)
    ref(variant).ref_get_x_b(m, n, _pi(_i32(rp)), _pi(_i32(ci)), _pv(v), _pv(x), _pv(b))
    return x, b


def ref_matrix_layer(m, rp, ci, v, variant="d64"):
    lv, par = c_int(), c_double()
    v = np.ascontiguousarray(v)
    ref(variant).ref_matrix_layer(m, m, len(ci), _pi(_i32(rp)), _pi(_i32(ci)), _pv(v), byref(lv), byref(par))
    return lv.value, par.value


def ref_transpose(m, n, nnz, rp, ci, v, variant="d64"):
    v = np.ascontiguousarray(v)
    cr = np.zeros(max(nnz, 1), np.int32)
    cp = np.zeros(n + 1, np.int32)
    cv = np.zeros(max(nnz, 1), v.dtype)
    ref(variant).ref_matrix_transposition(m, n, nnz, _pi(_i32(rp)), _pi(_i32(ci)), _pv(v), _pi(cr), _pi(cp), _pv(cv))
    return cr[:nnz], cp, cv[:nnz]


def ref_matrix_warp4(m, nnz, rp, ci, v, border, variant="d64"):
    wn = np.zeros(m + 1, np.int32)
    ln, wo, eo = c_int(), c_double(), c_double()
    v = np.ascontiguousarray(v)
    ref(variant).ref_matrix_warp4(m, m, nnz, _pi(_i32(rp)), _pi(_i32(ci)), _pv(v), border, byref(ln), _pi(wn),
                                  byref(wo), byref(eo))
    return ln.value, wn[:ln.value].copy(), wo.value, eo.value


def ref_csc_executor(cp, ri, cv, indeg, m, n, substitution, rhs, b, variant="d64"):
    cv = np.ascontiguousarray(cv)
    b = np.ascontiguousarray(b, dtype=cv.dtype)
    x = np.zeros(n * rhs, cv.dtype)
    ref(variant).ref_serial_executor(_pi(_i32(cp)), _pi(_i32(ri)), _pv(cv), _pi(_i32(indeg)), m, n, substitution,
                                     rhs, _pv(b), _pv(x))
    return x


def ref_analyser(csc_row, m, nnz, variant="d64"):
    deg = np.zeros(max(m, 1), np.int32)
    ref(variant).ref_serial_analyser(_pi(_i32(csc_row)), m, m, nnz, _pi(deg))
    return deg[:m]
