"""ctypes binding of libcapellini_sptrsv.so (the C ABI in include/capellini_sptrsv.h).

There is no fallback: if the shared library is missing or does not export a
declared symbol, loading raises NativeLibraryError.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_double, c_int, c_longlong, c_ulonglong, c_void_p, c_char_p, c_float

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcapellini_sptrsv.so")

CAP_OK = 0
CAP_ERR_NOT_SQUARE = -1
CAP_ERR_ARG = -2
CAP_ERR_ROWPTR = -3
CAP_ERR_UPPER = -4
CAP_ERR_NO_DIAG = -5
CAP_ERR_ZERO_DIAG = -6
CAP_ERR_HIP = -7
CAP_ERR_NO_DEVICE = -8
CAP_ERR_ALLOC = -9
CAP_ERR_UNSUPPORTED = -10
CAP_ERR_IO = -11
CAP_ERR_NOT_CONVERGED = -12
CAP_ERR_WATCHDOG = -13

CAP_F64, CAP_F32 = 0, 1
ALGO = {"auto": 0, "wrtfst": 1, "twopart": 2, "mix": 3, "csc": 4, "local": 5, "partition": 6, "level": 7}
ALGO_NAME = {v: k for k, v in ALGO.items()}
SYNC = {"auto": 0, "value": 1, "flag": 2}
ADJOINT = {"auto": 0, "scatter": 1, "gather": 2}
READINESS = {"auto": 0, "staged": 1, "in_place": 2}
READINESS_NAME = {v: k for k, v in READINESS.items()}
# fused: fma per term; reference: rounded product then rounded sum (the reference's serial executor)
ARITH = {"fused": 0, "reference": 1}
SYNC_NAME = {v: k for k, v in SYNC.items()}


class NativeLibraryError(RuntimeError):
    """The HIP library is not built or not loadable."""


class CapError(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        msg = _lib_strerror(code)
        super().__init__(f"{what} failed with {code} ({msg})")


class cap_opts(Structure):
    _fields_ = [("algo", c_int), ("sync", c_int), ("border", c_int), ("validate", c_int),
                ("host_inputs", c_int), ("spin_limit", c_int), ("blocks_per_cu", c_int),
                ("adjoint", c_int), ("lds_window", c_int), ("readiness", c_int),
                ("upper", c_int), ("unit_diag", c_int), ("row_threads", c_int), ("arith", c_int),
                ("count_polls", c_int), ("wait_limit_ms", c_int), ("rhs_per_lane", c_int),
                ("row_tiles", c_int)]


class cap_gen_params(Structure):
    _fields_ = [("kind", c_int), ("n", c_int), ("offdiag", c_int), ("band", c_int), ("levels", c_int),
                ("width", c_int), ("depth", c_int), ("p", c_double), ("seed", c_ulonglong)]


PI = POINTER(c_int)
PPI = POINTER(POINTER(c_int))
PD = POINTER(c_double)
PF = POINTER(c_float)
PPD = POINTER(POINTER(c_double))
PPF = POINTER(POINTER(c_float))

# symbol -> (restype, argtypes); the full exported surface of the header
SIGNATURES = {
    "cap_opts_default": (None, [POINTER(cap_opts)]),
    "cap_plan_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_void_p, c_void_p, c_void_p, c_int,
                                POINTER(cap_opts), c_void_p]),
    "cap_plan_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "cap_plan_status": (c_int, [c_void_p, c_void_p]),
    "cap_plan_reset": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "cap_plan_launch": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "cap_plan_info": (c_int, [c_void_p, PI, PI, PI, PI, PD, PD]),
    "cap_plan_poll_count": (c_int, [c_void_p, POINTER(c_ulonglong), c_void_p]),
    "cap_plan_rhs_block": (c_int, [c_void_p, c_int, PI, PI, PI, PI, PI, PI]),
    "cap_plan_tuning": (c_int, [c_void_p, PI, PI, PI, PI, PI, PI]),
    "cap_batch_info": (c_int, [c_void_p, PI, PI]),
    "cap_batch_tuning": (c_int, [c_void_p, PI, PI, PI, PI]),
    "cap_batch_partitioned": (c_int, [c_void_p, PI, c_int]),
    "cap_plan_destroy": (c_int, [c_void_p]),
    "cap_plan_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "cap_plan_solve_refined": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_double, c_int, PI, PD, c_void_p]),
    "cap_batch_create": (c_int, [POINTER(c_void_p), c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                 POINTER(cap_opts), c_void_p]),
    "cap_batch_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "cap_batch_solve_refined": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_int, PI, PD, c_void_p]),
    "cap_batch_destroy": (c_int, [c_void_p]),
    "sptrsv_csr_solve_d": (c_int, [c_int, c_int, PI, PI, PD, PD, PD, POINTER(cap_opts)]),
    "sptrsv_csr_solve_s": (c_int, [c_int, c_int, PI, PI, PF, PF, PF, POINTER(cap_opts)]),
    "sptrsv_csr_solve_noalloc_d": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    "spts_syncfree_hip_csr_wrt_d": (c_int, [PI, PI, PD, c_int, c_int, c_int, PD, PD, PD]),
    "spts_syncfree_hip_csr_wrt_s": (c_int, [PI, PI, PF, c_int, c_int, c_int, PD, PD, PD]),
    "spts_syncfree_hip_csr_algo_d": (c_int, [PI, PI, PD, c_int, c_int, c_int, c_int, PD, PD, PD]),
    "spts_syncfree_hip_csr_algo_s": (c_int, [PI, PI, PF, c_int, c_int, c_int, c_int, PD, PD, PD]),
    "spts_syncfree_hip_csr_d": (c_int, [PI, PI, PD, c_int, c_int, c_int, PD, PD, PD]),
    "spts_syncfree_hip_csr_s": (c_int, [PI, PI, PF, c_int, c_int, c_int, PD, PD, PD]),
    "spts_syncfree_hip_mix_d": (c_int, [PI, PI, PD, c_int, c_int, c_int, PD, PD, PD, PD, PD, PD]),
    "spts_syncfree_hip_mix_s": (c_int, [PI, PI, PF, c_int, c_int, c_int, PD, PD, PD, PD, PD, PD]),
    "sptrsv_syncfree_hip_csc_d": (c_int, [PI, PI, PD, c_int, c_int, c_int, PD, PD, PD, PD, PD, PD, PD]),
    "sptrsv_syncfree_hip_csc_s": (c_int, [PI, PI, PF, c_int, c_int, c_int, PF, PF, PF, PD, PD, PD, PD]),
    "cap_read_mtx_d": (c_int, [c_char_p, PI, PI, PI, PPI, PPI, PPD]),
    "cap_read_mtx_s": (c_int, [c_char_p, PI, PI, PI, PPI, PPI, PPF]),
    "cap_change2tran_d": (c_int, [c_int, c_int, PI, PI, PD, c_int, PI, PPI, PPI, PPD]),
    "cap_change2tran_s": (c_int, [c_int, c_int, PI, PI, PF, c_int, PI, PPI, PPI, PPF]),
    "cap_get_x_b_d": (None, [c_int, c_int, PI, PI, PD, PD, PD]),
    "cap_get_x_b_s": (None, [c_int, c_int, PI, PI, PF, PF, PF]),
    "cap_matrix_layer": (c_int, [c_int, PI, PI, PI, PD, PI]),
    "cap_matrix_warp4": (None, [c_int, c_int, PI, c_int, c_int, PI, PI, PD, PD]),
    "cap_transpose_d": (None, [c_int, c_int, c_int, PI, PI, PD, PI, PI, PD]),
    "cap_transpose_s": (None, [c_int, c_int, c_int, PI, PI, PF, PI, PI, PF]),
    "cap_validate_csr_d": (c_int, [c_int, c_int, PI, PI, PD, PI]),
    "cap_validate_csr_s": (c_int, [c_int, c_int, PI, PI, PF, PI]),
    "cap_generate_d": (c_int, [POINTER(cap_gen_params), PI, PI, PI, PD]),
    "cap_generate_s": (c_int, [POINTER(cap_gen_params), PI, PI, PI, PF]),
    "cap_uniform_d": (None, [c_ulonglong, c_longlong, c_double, c_double, PD]),
    "cap_serial_analyser": (c_int, [PI, c_int, c_int, c_int, PI]),
    "cap_serial_executor_d": (c_int, [PI, PI, PD, PI, c_int, c_int, c_int, c_int, PD, PD]),
    "cap_serial_executor_s": (c_int, [PI, PI, PF, PI, c_int, c_int, c_int, c_int, PF, PF]),
    "sptrsv_syncfree_serialref_d": (c_int, [PI, PI, PD, c_int, c_int, c_int, c_int, c_int, PD, PD, PD]),
    "sptrsv_syncfree_serialref_s": (c_int, [PI, PI, PF, c_int, c_int, c_int, c_int, c_int, PF, PF, PF]),
    "cap_serial_csr_d": (c_int, [c_int, c_int, PI, PI, PD, c_int, PD, PD, c_int]),
    "cap_serial_csr_s": (c_int, [c_int, c_int, PI, PI, PF, c_int, PF, PF, c_int]),
    "cap_datasize": (c_double, [c_int, c_int, c_int, c_int, c_int]),
    "cap_device_count": (c_int, []),
    "cap_strerror": (c_char_p, [c_int]),
    "cap_version": (c_char_p, []),
    "cap_free": (None, [c_void_p]),
}

_LIB = None


def lib() -> ctypes.CDLL:
    """Load the library once; raise NativeLibraryError if it is absent or incomplete."""
    global _LIB
    if _LIB is not None:
        return _LIB
    # One HIP runtime per process: torch ships its own libamdhip64 (same
    # SONAME, different file).  Loading torch first makes this library bind
    # to torch's runtime; loading this library first would pull in the
    # system runtime and torch would then load a second one, after which the
    # two cannot share the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python capellinisptrsv_amd/build.py`")
    try:
        h = ctypes.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - loader failure
        raise NativeLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    missing = []
    for name, (res, args) in SIGNATURES.items():
        try:
            fn = getattr(h, name)
        except AttributeError:
            missing.append(name)
            continue
        fn.restype = res
        fn.argtypes = args
    if missing:
        raise NativeLibraryError(f"{LIB_PATH} lacks symbols: {missing}")
    _LIB = h
    return h


def _lib_strerror(code: int) -> str:
    try:
        return lib().cap_strerror(code).decode()
    except NativeLibraryError:
        return "?"


def check(code: int, what: str) -> None:
    if code != CAP_OK:
        raise CapError(code, what)


def default_opts(**kw) -> cap_opts:
    o = cap_opts()
    lib().cap_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        if k == "algo" and isinstance(v, str):
            v = ALGO[v]
        if k == "sync" and isinstance(v, str):
            v = SYNC[v]
        if k == "adjoint" and isinstance(v, str):
            v = ADJOINT[v]
        if k == "readiness" and isinstance(v, str):
            v = READINESS[v]
        if k == "arith" and isinstance(v, str):
            v = ARITH[v]
        setattr(o, k, int(v))
    return o
