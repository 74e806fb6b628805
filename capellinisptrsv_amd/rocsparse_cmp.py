"""rocSPARSE `csrsv` as a comparator for the benchmark (context only).

The reference compares its kernels with cuSPARSE's level-scheduled and
level-free triangular solves (CUDA/cuSP/spts_cuSPARSE.h:109-154,
CUDA/cuSP-layer/spts_cuSPARSE.h:122-186).  On MI355X the vendor routine
is rocSPARSE's `csrsv`: an analysis pass once per matrix, then solves.
This module binds the C API with ctypes; it loads the librocsparse that
ships with torch, so the process keeps a single HIP runtime.  Nothing in
the library's solve path uses it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_double, c_float, c_int, c_size_t, c_void_p

import torch

_OP_NONE = 111  # rocsparse_operation_none
_FILL_LOWER = 0  # rocsparse_fill_mode_lower
_DIAG_NON_UNIT = 0  # rocsparse_diag_type_non_unit
_ANALYSIS_REUSE = 0  # rocsparse_analysis_policy_reuse
_SOLVE_AUTO = 0  # rocsparse_solve_policy_auto

_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(os.path.dirname(torch.__file__), "lib", "librocsparse.so")
        _lib = ctypes.CDLL(path)
        for pre, ct in (("d", c_double), ("s", c_float)):
            getattr(_lib, f"rocsparse_{pre}csrsv_buffer_size").argtypes = [
                c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, POINTER(c_size_t)]
            getattr(_lib, f"rocsparse_{pre}csrsv_analysis").argtypes = [
                c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                c_void_p]
            getattr(_lib, f"rocsparse_{pre}csrsv_solve").argtypes = [
                c_void_p, c_int, c_int, c_int, POINTER(ct), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                c_void_p, c_void_p, c_int, c_void_p]
        _lib.rocsparse_set_stream.argtypes = [c_void_p, c_void_p]
    return _lib


def _check(status: int, what: str):
    if status != 0:
        raise RuntimeError(f"{what} returned rocsparse_status {status}")


class Csrsv:
    """y = L^{-1} x with rocSPARSE (lower, non-unit diagonal, sorted CSR)."""

    def __init__(self, row_ptr: torch.Tensor, col_idx: torch.Tensor, values: torch.Tensor):
        lib = _load()
        self.rp, self.ci, self.val = row_ptr, col_idx, values
        self.n = row_ptr.numel() - 1
        self.nnz = col_idx.numel()
        self.pre = {torch.float64: "d", torch.float32: "s"}[values.dtype]
        self.ct = c_double if self.pre == "d" else c_float
        self.handle, self.descr, self.info = c_void_p(), c_void_p(), c_void_p()
        _check(lib.rocsparse_create_handle(byref(self.handle)), "rocsparse_create_handle")
        _check(lib.rocsparse_set_stream(self.handle, c_void_p(torch.cuda.current_stream().cuda_stream)),
               "rocsparse_set_stream")
        _check(lib.rocsparse_create_mat_descr(byref(self.descr)), "rocsparse_create_mat_descr")
        _check(lib.rocsparse_set_mat_fill_mode(self.descr, _FILL_LOWER), "rocsparse_set_mat_fill_mode")
        _check(lib.rocsparse_set_mat_diag_type(self.descr, _DIAG_NON_UNIT), "rocsparse_set_mat_diag_type")
        _check(lib.rocsparse_create_mat_info(byref(self.info)), "rocsparse_create_mat_info")
        size = c_size_t()
        _check(getattr(lib, f"rocsparse_{self.pre}csrsv_buffer_size")(
            self.handle, _OP_NONE, self.n, self.nnz, self.descr, c_void_p(values.data_ptr()),
            c_void_p(row_ptr.data_ptr()), c_void_p(col_idx.data_ptr()), self.info, byref(size)),
            "csrsv_buffer_size")
        self.buffer = torch.empty(max(1, size.value), dtype=torch.uint8, device=values.device)
        self._alpha = self.ct(1.0)

    def analysis(self):
        _check(getattr(_load(), f"rocsparse_{self.pre}csrsv_analysis")(
            self.handle, _OP_NONE, self.n, self.nnz, self.descr, c_void_p(self.val.data_ptr()),
            c_void_p(self.rp.data_ptr()), c_void_p(self.ci.data_ptr()), self.info, _ANALYSIS_REUSE, _SOLVE_AUTO,
            c_void_p(self.buffer.data_ptr())), "csrsv_analysis")

    def solve(self, b: torch.Tensor, out: torch.Tensor):
        _check(getattr(_load(), f"rocsparse_{self.pre}csrsv_solve")(
            self.handle, _OP_NONE, self.n, self.nnz, byref(self._alpha), self.descr, c_void_p(self.val.data_ptr()),
            c_void_p(self.rp.data_ptr()), c_void_p(self.ci.data_ptr()), self.info, c_void_p(b.data_ptr()),
            c_void_p(out.data_ptr()), _SOLVE_AUTO, c_void_p(self.buffer.data_ptr())), "csrsv_solve")
        return out

    def close(self):
        lib = _load()
        if self.info:
            lib.rocsparse_destroy_mat_info(self.info)
            self.info = c_void_p()
        if self.descr:
            lib.rocsparse_destroy_mat_descr(self.descr)
            self.descr = c_void_p()
        if self.handle:
            lib.rocsparse_destroy_handle(self.handle)
            self.handle = c_void_p()
