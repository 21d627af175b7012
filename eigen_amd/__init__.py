"""eigen_amd -- MI355X (gfx950) dense factor-and-solve: Eigen's GEMM,
PartialPivLU, LLT and triangular solves as hand-written HIP kernels behind a
C ABI (include/eigen_amd.h), plus mixed-precision LU with iterative
refinement and batched small solves.

The HIP library must be built (``make -C eigen_amd/csrc``); there is no CPU
fallback.
"""
from ._lib import EamdError, declared_symbols, lib  # noqa: F401
from .dense import (LDLT, LLT, InvalidInput, NoConvergence, NumericalIssue,  # noqa: F401
                    PartialPivLU, Success, backward_error, colmajor, gemm, handle,
                    rank_update, triangular_solve)

__version__ = "0.1.0"
