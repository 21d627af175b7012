"""Batched small solves over packed variable-size systems (device tensors).

Per system this is ``LLT<MatrixXf>(A_i).solve(b_i)`` (LLT) or
``PartialPivLU<MatrixXd>(A_i).solve(B_i)`` (LU with FP64 refinement); the
systems are packed column-major back to back, located by an offset array,
and solved in one captured launch sequence.
"""
import ctypes as C

import torch

from ._lib import check, lib
from .dense import handle


def _p(t):
    return C.c_void_p(t.data_ptr())


class BatchedPlan:
    """Size-bucketing plan for a packed batch (sizes n: int32 device tensor)."""

    def __init__(self, n, max_n=32):
        self.count = int(n.numel())
        self.ptr = C.c_void_p()
        check("eamd_batched_plan_create",
              lib().eamd_batched_plan_create(handle(), self.count, _p(n), max_n,
                                             C.byref(self.ptr)))

    def lanes(self, n):
        return lib().eamd_batched_plan_lanes(self.ptr, n)

    def solve_llt(self, offA, A, offb, b, info):
        """b <- A_i^{-1} b_i per system (in place); info[i] = failing column or -1."""
        check("eamd_spotrf_solve_batched_planned",
              lib().eamd_spotrf_solve_batched_planned(handle(), self.ptr, _p(offA), _p(A),
                                                      _p(offb), _p(b), _p(info)))

    def __del__(self):
        try:
            if self.ptr:
                lib().eamd_batched_plan_destroy(self.ptr)
        except Exception:
            pass


class BatchedLuPlan:
    """Plan of a batched LU + FP64 refinement solve (sizes n: int32 device
    tensor, 1..256; nrhs right-hand sides per system)."""

    def __init__(self, n, nrhs):
        self.count = int(n.numel())
        self.nrhs = int(nrhs)
        self.ptr = C.c_void_p()
        check("eamd_batched_lu_plan_create",
              lib().eamd_batched_lu_plan_create(handle(), self.count, _p(n), self.nrhs,
                                                C.byref(self.ptr)))

    def threads(self, n):
        """Workgroup size the autotuner chose for systems of size n."""
        return lib().eamd_batched_plan_lanes(self.ptr, n)

    def solve(self, offA, n, A, offB, B, X, info, iters=None):
        check("eamd_dgesv_ir_batched_planned",
              lib().eamd_dgesv_ir_batched_planned(
                  handle(), self.ptr, _p(offA), _p(n), _p(A), _p(offB), _p(B), _p(X), _p(info),
                  _p(iters) if iters is not None else None))

    def __del__(self):
        try:
            if self.ptr:
                lib().eamd_batched_plan_destroy(self.ptr)
        except Exception:
            pass


def spotrf_solve_batched(offA, n, A, offb, b):
    """One-shot batched LLT solve: b overwritten with x; returns info."""
    info = torch.empty(n.numel(), dtype=torch.int32, device=n.device)
    check("eamd_spotrf_solve_batched",
          lib().eamd_spotrf_solve_batched(handle(), n.numel(), _p(offA), _p(n), _p(A), _p(offb),
                                          _p(b), _p(info)))
    return info


def llt_rank_update_batched(offL, n, L, offv, v, sigma=None, upper=False):
    """LLT::rankUpdate (LLT.h:459-472) for packed factors: L (FP64, factor i
    column-major n[i] x n[i] at offL[i]) updated in place so that
    L_i L_i^T = A_i + sigma_i v_i v_i^T; returns info (-1, the failing column
    of a downdate, or -2 for n outside 0..256)."""
    info = torch.empty(n.numel(), dtype=torch.int32, device=n.device)
    check("eamd_dllt_rank_update_batched",
          lib().eamd_dllt_rank_update_batched(handle(), n.numel(), _p(offL), _p(n), _p(L),
                                              _p(offv), _p(v),
                                              _p(sigma) if sigma is not None else None,
                                              1 if upper else 0, _p(info)))
    return info


def ldlt_batched(offA, n, A, offT=None, max_n=None):
    """LDLT<MatrixXd, Lower>::compute (LDLT.h:491-522) for packed symmetric
    matrices: A (FP64, matrix i column-major n[i] x n[i] at offA[i]) has its
    lower triangle replaced by matrixLDLT()'s.  Returns (T, sign, info):
    transpositions (int32, system i's n[i] entries at offT[i]; offT defaults
    to the prefix sums of n), sign per system (0 PositiveSemiDef,
    1 NegativeSemiDef, 2 ZeroSign, 3 Indefinite) and info (0 Success,
    1 NumericalIssue, -2 for n outside 0..max_n).  max_n (default: the
    largest n, read once on the host) is at most 64."""
    dev = n.device
    count = n.numel()
    if max_n is None:
        max_n = int(n.max()) if count else 0
    if offT is None:
        offT = (torch.cumsum(n.to(torch.int64), 0) - n.to(torch.int64)).contiguous()
    # the furthest end of any system's slots (offT need not be ascending)
    total = int((offT + n.to(torch.int64).clamp(min=0)).max()) if count else 0
    T = torch.zeros(max(total, 1), dtype=torch.int32, device=dev)
    sign = torch.empty(count, dtype=torch.int32, device=dev)
    info = torch.empty(count, dtype=torch.int32, device=dev)
    check("eamd_dldlt_batched",
          lib().eamd_dldlt_batched(handle(), count, _p(offA), _p(n), int(max_n), _p(A), _p(offT),
                                   _p(T), _p(sign), _p(info)))
    return T, offT, sign, info


def ldlt_solve_batched(offA, n, A, offT, T, offb, b, max_n=None):
    """LDLT::solve (LDLT.h:559-593) with ldlt_batched's factors: b (FP64,
    system i's right-hand side at offb[i]) overwritten with the solution;
    max_n as for ldlt_batched."""
    if max_n is None:
        max_n = int(n.max()) if n.numel() else 0
    check("eamd_dldlt_solve_batched",
          lib().eamd_dldlt_solve_batched(handle(), n.numel(), _p(offA), _p(n), int(max_n), _p(A),
                                         _p(offT), _p(T), _p(offb), _p(b)))
    return b


def ldlt_rank_update_batched(offA, n, A, offT, T, offw, w, sigma=None):
    """LDLT::rankUpdate (LDLT.h:403-449, 529-555) for ldlt_batched's factors:
    factor i (A at offA[i], transpositions at offT[i]) updated in place to
    the factor of A_i + sigma_i w_i w_i^T, w_i (FP64, n[i] values at offw[i])
    read only; sigma (FP64 per factor) defaults to 1.  The transpositions
    and sign are unchanged, as in the reference.  Returns info (0, or -2 for
    n outside 0..256)."""
    info = torch.empty(n.numel(), dtype=torch.int32, device=n.device)
    check("eamd_dldlt_rank_update_batched",
          lib().eamd_dldlt_rank_update_batched(handle(), n.numel(), _p(offA), _p(n), _p(A),
                                               _p(offT), _p(T), _p(offw), _p(w),
                                               _p(sigma) if sigma is not None else None,
                                               _p(info)))
    return info


def dgesv_ir_batched(offA, n, A, offB, nrhs, B):
    """Batched FP32-LU + FP64-refinement solve; returns (X, info, iters)."""
    X = torch.empty_like(B)
    info = torch.empty(n.numel(), dtype=torch.int32, device=n.device)
    iters = torch.empty(n.numel(), dtype=torch.int32, device=n.device)
    check("eamd_dgesv_ir_batched",
          lib().eamd_dgesv_ir_batched(handle(), n.numel(), _p(offA), _p(n), _p(A), _p(offB),
                                      nrhs, _p(B), _p(X), _p(info), _p(iters)))
    return X, info, iters


def capture_llt(plan, offA, A, offb, b, info):
    """A torch CUDA graph replaying the planned batched LLT on these buffers."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.solve_llt(offA, A, offb, b, info)  # warm the kernels outside capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    # capture on the stream the library handle is already bound to, so binding
    # inside the capture is a no-op (no synchronisation while capturing)
    with torch.cuda.graph(g, stream=s):
        plan.solve_llt(offA, A, offb, b, info)
    return g
