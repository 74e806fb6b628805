"""torch-facing wrappers: device-resident plans, batches and the
differentiable solve.  torch provides device memory and streams only; every
solve, adjoint and refinement runs in the HIP kernels of
libcapellini_sptrsv.so through its C ABI.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_double, c_int, c_void_p

import numpy as np
import torch

from . import _native as N

__all__ = ["Plan", "Batch", "SpTRSV", "sptrsv", "TriangularSolve"]


def _stream(stream=None) -> c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return c_void_p(s.cuda_stream)


def _ptr(t: torch.Tensor | None) -> c_void_p:
    return c_void_p(t.data_ptr()) if t is not None else c_void_p(0)


def _vt(dtype) -> int:
    if dtype == torch.float64:
        return N.CAP_F64
    if dtype == torch.float32:
        return N.CAP_F32
    raise TypeError(f"values must be float64 or float32, got {dtype}")


class Plan:
    """One lower-triangular CSR matrix resident on the GPU.

    row_ptr/col_idx: int32 CUDA tensors; values: float64/float32 CUDA tensor.
    The plan keeps references to the three tensors (their memory is used in
    place).  algo: auto|wrtfst|twopart|mix|csc|local|partition|level; sync:
    auto|value|flag.
    row_threads: 128 or 256 threads per workgroup of the thread-per-row
    kernels (0: tuned with algo auto, else 256).
    upper=True: upper triangular, diagonal first in each row; unit_diagonal=
    True: no diagonal stored (taken as 1).  Both convert the matrix into a
    plan-owned copy; such plans solve, but have no adjoint or refinement.
    spin_limit / wait_limit_ms: a dependency wait gives up after that many
    polls / that much wall time without progress (0: no poll bound, 10 s).
    arith: "fused" (one fused multiply-add per term) or "reference" (a
    rounded product then a rounded sum per term: bit-identical to the
    reference's serial CPU executor; not with csc or partition).
    row_tiles (value protocol): each ticket's col_idx/val range is copied to
    LDS with coalesced loads first (algo auto times it).
    rhs_per_lane (algo wrtfst, value protocol): solves of several right-hand
    sides run in 1024-thread workgroups with that many per lane (1, 2, 4;
    8 for FP32); 0 leaves the choice to the workgroup size above (or to the
    autotuner with algo auto).
    """

    def __init__(self, row_ptr: torch.Tensor, col_idx: torch.Tensor, values: torch.Tensor, algo="auto",
                 sync="auto", border: int = 0, validate: bool = True, spin_limit: int = 0,
                 blocks_per_cu: int = 0, adjoint="auto", lds_window: bool = False, readiness="auto",
                 upper: bool = False, unit_diagonal: bool = False, row_threads: int = 0, arith="fused",
                 wait_limit_ms: int = 0, count_polls: bool = False, rhs_per_lane: int = 0, row_tiles: bool = False,
                 stream=None):
        for t, dt in ((row_ptr, torch.int32), (col_idx, torch.int32)):
            if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
                raise TypeError("row_ptr/col_idx must be contiguous int32 CUDA tensors")
        if not values.is_cuda or not values.is_contiguous():
            raise TypeError("values must be a contiguous CUDA tensor")
        self.row_ptr, self.col_idx, self.values = row_ptr, col_idx, values
        self.n = row_ptr.numel() - 1
        self.nnz = col_idx.numel()
        self.dtype = values.dtype
        opts = N.default_opts(algo=algo, sync=sync, border=border, validate=int(validate), host_inputs=0,
                              spin_limit=spin_limit, blocks_per_cu=blocks_per_cu, adjoint=adjoint,
                              lds_window=int(lds_window), readiness=readiness, upper=int(upper),
                              unit_diag=int(unit_diagonal), row_threads=row_threads, arith=arith,
                              wait_limit_ms=wait_limit_ms, count_polls=int(count_polls),
                              rhs_per_lane=rhs_per_lane, row_tiles=int(row_tiles))
        h = c_void_p()
        N.check(N.lib().cap_plan_create(byref(h), self.n, self.nnz, _ptr(row_ptr), _ptr(col_idx), _ptr(values),
                                        _vt(values.dtype), byref(opts), _stream(stream)), "cap_plan_create")
        self._h = h

    # -- queries
    def info(self) -> dict:
        a, s, b, ln = c_int(), c_int(), c_int(), c_int()
        wo, eo = c_double(), c_double()
        N.check(N.lib().cap_plan_info(self._h, byref(a), byref(s), byref(b), byref(ln), byref(wo), byref(eo)),
                "cap_plan_info")
        bpc, lds, adj, rdy, thr, tile = c_int(), c_int(), c_int(), c_int(), c_int(), c_int()
        N.check(N.lib().cap_plan_tuning(self._h, byref(bpc), byref(lds), byref(adj), byref(rdy), byref(thr),
                                        byref(tile)), "cap_plan_tuning")
        return dict(algo=N.ALGO_NAME[a.value], sync=N.SYNC_NAME[s.value], border=b.value, mix_len=ln.value,
                    warp_occupy=wo.value, element_occupy=eo.value, blocks_per_cu=bpc.value,
                    lds_window=bool(lds.value), adjoint={1: "scatter", 2: "gather"}.get(adj.value, "?"),
                    readiness=N.READINESS_NAME.get(rdy.value, "?"), row_threads=thr.value,
                    row_tiles=bool(tile.value))

    def poll_count(self, stream=None) -> int:
        """Poll loads issued by the thread-per-row solves since the last call
        (plans created with count_polls=True); synchronizes and clears."""
        from ctypes import c_ulonglong
        v = c_ulonglong()
        N.check(N.lib().cap_plan_poll_count(self._h, byref(v), _stream(stream)), "cap_plan_poll_count")
        return v.value

    def rhs_block(self, k: int):
        """The RHS block the autotuner picked for k right-hand sides (after a
        solve with k): dict(blocks_per_cu, row_threads, rhs_per_wave,
        rhs_per_lane, readiness, row_tiles), or None."""
        bpc, thr, kr, rpl, rdy, tile = c_int(), c_int(), c_int(), c_int(), c_int(), c_int()
        rc = N.lib().cap_plan_rhs_block(self._h, k, byref(bpc), byref(thr), byref(kr), byref(rpl), byref(rdy),
                                        byref(tile))
        if rc == N.CAP_ERR_UNSUPPORTED:
            return None
        N.check(rc, "cap_plan_rhs_block")
        return dict(blocks_per_cu=bpc.value, row_threads=thr.value, rhs_per_wave=kr.value, rhs_per_lane=rpl.value,
                    readiness=N.READINESS_NAME.get(rdy.value, "?"), row_tiles=bool(tile.value))

    def status(self, stream=None) -> int:
        """Synchronizes, returns and clears the watchdog code (0 = every wait completed)."""
        return N.lib().cap_plan_status(self._h, _stream(stream))

    # -- solves
    def solve(self, b: torch.Tensor, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        """x = L^{-1} b; b is (n,) or (n, k) with k interleaved right-hand sides (row-major)."""
        if b.dtype != self.dtype or not b.is_cuda:
            raise TypeError("b must be a CUDA tensor of the plan's dtype")
        b = b.contiguous()
        k = 1 if b.dim() == 1 else b.shape[1]
        if b.shape[0] != self.n:
            raise ValueError("b has the wrong number of rows")
        x = out if out is not None else torch.empty_like(b)
        N.check(N.lib().cap_plan_solve(self._h, _ptr(b), _ptr(x), k, _stream(stream)), "cap_plan_solve")
        return x

    def reset(self, x: torch.Tensor, stream=None):
        """First half of solve(): per-solve reset of x / tickets (asynchronous)."""
        k = 1 if x.dim() == 1 else x.shape[1]
        N.check(N.lib().cap_plan_reset(self._h, _ptr(x), k, _stream(stream)), "cap_plan_reset")

    def launch(self, b: torch.Tensor, x: torch.Tensor, stream=None):
        """Second half of solve(): the solve kernel alone (after reset(x))."""
        k = 1 if b.dim() == 1 else b.shape[1]
        N.check(N.lib().cap_plan_launch(self._h, _ptr(b), _ptr(x), k, _stream(stream)), "cap_plan_launch")
        return x

    def adjoint(self, g: torch.Tensor, x: torch.Tensor, want_grad_val: bool = True, stream=None):
        """(grad_b, grad_val) of a loss with dL/dx = g at the solution x (FP64)."""
        if self.dtype != torch.float64:
            raise TypeError("the adjoint is FP64 only")
        g = g.contiguous()
        gb = torch.empty_like(g)
        gv = torch.empty_like(self.values) if want_grad_val else None
        N.check(N.lib().cap_plan_adjoint(self._h, _ptr(g), _ptr(x.contiguous()), _ptr(gb), _ptr(gv),
                                         _stream(stream)), "cap_plan_adjoint")
        return gb, gv

    def solve_refined(self, b: torch.Tensor, tol: float = 1e-12, max_iter: int = 8, out=None, stream=None):
        """FP32 solves inside FP64 iterative refinement.  Returns (x, iterations, last relative correction)."""
        if self.dtype != torch.float64:
            raise TypeError("refinement needs an FP64 plan")
        b = b.contiguous()
        k = 1 if b.dim() == 1 else b.shape[1]
        x = out if out is not None else torch.empty_like(b)
        it, rel = c_int(), c_double()
        rc = N.lib().cap_plan_solve_refined(self._h, _ptr(b), _ptr(x), k, tol, max_iter, byref(it), byref(rel),
                                            _stream(stream))
        N.check(rc, "cap_plan_solve_refined")
        return x, it.value, rel.value

    def close(self):
        if getattr(self, "_h", None):
            N.lib().cap_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many matrices (concatenated, local column indices) with k interleaved RHS."""

    def __init__(self, row_off: torch.Tensor, row_ptr: torch.Tensor, col_idx: torch.Tensor, values: torch.Tensor,
                 k: int, sync="value", validate: bool = True, algo="auto", blocks_per_cu: int = 0,
                 arith="fused", stream=None):
        """algo: auto (both timed at creation) | wrtfst (level-scheduled
        thread-per-row over all matrices) | local (one workgroup per matrix).
        blocks_per_cu: persistent workgroups per compute unit (0: chosen
        at creation).  arith: "fused" or "reference" (as for Plan)."""
        self.row_off, self.row_ptr, self.col_idx, self.values = row_off, row_ptr, col_idx, values
        self.k = k
        self.n = row_ptr.numel() - 1
        self.dtype = values.dtype
        opts = N.default_opts(algo=algo, sync=sync, validate=int(validate), host_inputs=0,
                              blocks_per_cu=blocks_per_cu, arith=arith)
        h = c_void_p()
        N.check(N.lib().cap_batch_create(byref(h), row_off.numel() - 1, _ptr(row_off), _ptr(row_ptr),
                                         _ptr(col_idx), _ptr(values), _vt(values.dtype), k, byref(opts),
                                         _stream(stream)), "cap_batch_create")
        self._h = h
        a, bpc = c_int(), c_int()
        N.check(N.lib().cap_batch_info(h, byref(a), byref(bpc)), "cap_batch_info")
        self.algo = N.ALGO_NAME[a.value]
        self.blocks_per_cu = bpc.value
        thr, rpl, npart, tile = c_int(), c_int(), c_int(), c_int()
        N.check(N.lib().cap_batch_tuning(h, byref(thr), byref(rpl), byref(npart), byref(tile)), "cap_batch_tuning")
        self.row_threads, self.rhs_per_lane, self.partitioned = thr.value, rpl.value, npart.value
        self.row_tiles = bool(tile.value)
        rows = (c_int * max(1, npart.value))()
        got = N.lib().cap_batch_partitioned(h, rows, npart.value)
        if got < 0:
            N.check(got, "cap_batch_partitioned")
        off = row_off.cpu().numpy() if row_off.is_cuda else row_off.numpy()
        # the matrices solved by partitioned plans (hybrid batch), by index
        self.partitioned_matrices = sorted(int(np.searchsorted(off, r, side="right") - 1) for r in rows[:got])

    def solve(self, b: torch.Tensor, out=None, stream=None) -> torch.Tensor:
        b = b.contiguous()
        x = out if out is not None else torch.empty_like(b)
        N.check(N.lib().cap_batch_solve(self._h, _ptr(b), _ptr(x), _stream(stream)), "cap_batch_solve")
        return x

    def solve_refined(self, b: torch.Tensor, tol: float = 1e-12, max_iter: int = 8, out=None, stream=None):
        b = b.contiguous()
        x = out if out is not None else torch.empty_like(b)
        it, rel = c_int(), c_double()
        N.check(N.lib().cap_batch_solve_refined(self._h, _ptr(b), _ptr(x), tol, max_iter, byref(it), byref(rel),
                                                _stream(stream)), "cap_batch_solve_refined")
        return x, it.value, rel.value

    def close(self):
        if getattr(self, "_h", None):
            N.lib().cap_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SpTRSV(torch.autograd.Function):
    """x = L^{-1} b with L given by (plan.row_ptr, plan.col_idx, values).

    backward: grad_b = L^{-T} grad_x (column-driven sync-free transposed
    solve), grad_values[k] = -grad_b[row_k] * x[col_k].
    """

    @staticmethod
    def forward(ctx, values: torch.Tensor, b: torch.Tensor, plan: Plan):
        if values.data_ptr() != plan.values.data_ptr():
            raise ValueError("values must be the tensor the plan was built on")
        x = plan.solve(b)
        ctx.plan = plan
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        (x,) = ctx.saved_tensors
        gb, gv = ctx.plan.adjoint(grad_x.contiguous(), x, want_grad_val=ctx.needs_input_grad[0])
        return gv, gb, None


def sptrsv(values: torch.Tensor, b: torch.Tensor, plan: Plan) -> torch.Tensor:
    return SpTRSV.apply(values, b, plan)


class TriangularSolve(torch.nn.Module):
    """Module holding the sparsity pattern and a learnable value vector."""

    def __init__(self, row_ptr, col_idx, values, algo="auto"):
        super().__init__()
        self.register_buffer("row_ptr", row_ptr)
        self.register_buffer("col_idx", col_idx)
        self.values = torch.nn.Parameter(values)
        self.plan = Plan(self.row_ptr, self.col_idx, self.values.data, algo=algo)

    def forward(self, b):
        return SpTRSV.apply(self.values, b, self.plan)
