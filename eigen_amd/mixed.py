"""Mixed-precision LU with iterative refinement (no counterpart in the
reference; it extends the PartialPivLU solve surface).

``gesv_ir(A, B, precision="fp16x3")`` factors a low-precision copy of A
(FP32 panels, trailing update on FP16/BF16 matrix cores with per-block
power-of-two scaling and FP32 accumulation), then refines in FP64 on the
device until the normwise backward error reaches ``tol`` (1e-14 default),
falling back to LU-preconditioned GMRES when the classic iteration stalls.
"""
import ctypes as C

import torch

from ._lib import IrOpts, IrReport, check, lib
from .dense import _ld, _p, colmajor, handle

PRECISIONS = {"fp64": 0, "fp32": 1, "fp16": 2, "bf16": 3, "fp16x3": 4}


def ir_options(precision="fp16x3", tol=1e-14, max_iter=30, gmres=1, gmres_restart=30,
               gmres_max_iter=300, nb=0):
    o = IrOpts()
    lib().eamd_ir_opts_default(C.byref(o))
    o.factor_precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    o.tol = tol
    o.max_iter = max_iter
    o.gmres = gmres
    o.gmres_restart = gmres_restart
    o.gmres_max_iter = gmres_max_iter
    o.nb = nb
    return o


def report_dict(r):
    return {k: getattr(r, k) for k, _ in IrReport._fields_}


def gesv_ir(A, B, precision="fp16x3", check_converged=True, **kw):
    """Solve A X = B; returns (X, report dict).  A, B: CUDA tensors (FP64)."""
    A = colmajor(A, torch.float64)
    vec = B.dim() == 1
    B = colmajor(B, torch.float64)
    n = A.shape[0]
    X = colmajor(torch.zeros(n, B.shape[1], dtype=torch.float64, device=A.device))
    o = ir_options(precision, **kw)
    rep = IrReport()
    st = lib().eamd_dgesv_ir(handle(), n, B.shape[1], _p(A), _ld(A), _p(B), _ld(B), _p(X),
                             _ld(X), C.byref(o), C.byref(rep))
    if st != -1003 or check_converged:
        check("eamd_dgesv_ir", st)
    return (X.reshape(-1) if vec else X), report_dict(rep)
