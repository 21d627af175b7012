"""Loader for the HIP library ``libeigen_amd.so`` and its C ABI.

The library is built in-tree (``make -C eigen_amd/csrc``).  There is no
fallback: if it is missing, importing a compute entry point raises.
"""
import ctypes as C
import os
import re
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeigen_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "eigen_amd.h")

_lib = None
_lock = threading.Lock()

c_int64 = C.c_int64
c_int = C.c_int
c_dbl = C.c_double
c_char = C.c_char
vp = C.c_void_p


class IrOpts(C.Structure):
    _fields_ = [("factor_precision", c_int), ("max_iter", c_int), ("tol", c_dbl),
                ("gmres", c_int), ("gmres_restart", c_int), ("gmres_max_iter", c_int),
                ("nb", c_int)]


class IrReport(C.Structure):
    _fields_ = [("ir_iterations", c_int), ("gmres_iterations", c_int), ("converged", c_int),
                ("used_gmres", c_int), ("backward_error", c_dbl), ("factor_ms", c_dbl),
                ("refine_ms", c_dbl), ("first_zero_pivot", c_int64)]


_SIGS = {
    "eamd_version": ([], c_int),
    "eamd_create": ([C.POINTER(vp)], c_int),
    "eamd_destroy": ([vp], c_int),
    "eamd_set_stream": ([vp, vp], c_int),
    "eamd_get_stream": ([vp], vp),
    "eamd_synchronize": ([vp], c_int),
    "eamd_last_error": ([], C.c_char_p),
    "eamd_release_workspace": ([vp], c_int),
    "eamd_profile_enable": ([vp, c_int], c_int),
    "eamd_profile_read": ([vp, c_int, C.POINTER(c_dbl), C.POINTER(c_dbl), C.POINTER(c_dbl),
                           C.POINTER(c_int64)], c_int),
    "eamd_dgemm": ([vp, c_char, c_char, c_int64, c_int64, c_int64, c_dbl, vp, c_int64, vp,
                    c_int64, c_dbl, vp, c_int64], c_int),
    "eamd_dtrsm": ([vp, c_char, c_char, c_char, c_char, c_int64, c_int64, c_dbl, vp, c_int64,
                    vp, c_int64], c_int),
    "eamd_dsyrk": ([vp, c_char, c_char, c_int64, c_int64, c_dbl, vp, c_int64, c_dbl, vp,
                    c_int64], c_int),
    "eamd_dgetrf": ([vp, c_int64, c_int64, vp, c_int64, vp, C.POINTER(C.c_int32),
                     C.POINTER(c_int64)], c_int),
    "eamd_dgetrs": ([vp, c_char, c_int64, c_int64, vp, c_int64, vp, vp, c_int64], c_int),
    "eamd_transpositions_to_permutation": ([vp, c_int64, vp, vp], c_int),
    "eamd_dpotrf": ([vp, c_char, c_int64, vp, c_int64, C.POINTER(c_int64)], c_int),
    "eamd_dpotrs": ([vp, c_char, c_int64, c_int64, vp, c_int64, vp, c_int64], c_int),
    "eamd_sgetrf": ([vp, c_int64, c_int64, vp, c_int64, vp, C.POINTER(C.c_int32),
                     C.POINTER(c_int64)], c_int),
    "eamd_sgetrs": ([vp, c_char, c_int64, c_int64, vp, c_int64, vp, vp, c_int64], c_int),
    "eamd_spotrf": ([vp, c_char, c_int64, vp, c_int64, C.POINTER(c_int64)], c_int),
    "eamd_spotrs": ([vp, c_char, c_int64, c_int64, vp, c_int64, vp, c_int64], c_int),
    "eamd_ir_opts_default": ([C.POINTER(IrOpts)], None),
    "eamd_dgesv_ir": ([vp, c_int64, c_int64, vp, c_int64, vp, c_int64, vp, c_int64,
                       C.POINTER(IrOpts), C.POINTER(IrReport)], c_int),
    "eamd_dbackward_error": ([vp, c_int64, c_int64, vp, c_int64, vp, c_int64, vp, c_int64,
                              C.POINTER(c_dbl)], c_int),
    "eamd_spotrf_solve_batched": ([vp, c_int64, vp, vp, vp, vp, vp, vp], c_int),
    "eamd_dldlt_rank_update_batched": ([vp, c_int64, vp, vp, vp, vp, vp, vp, vp, vp, vp], c_int),
    "eamd_dllt_rank_update_batched": ([vp, c_int64, vp, vp, vp, vp, vp, vp, C.c_int32, vp], c_int),
    "eamd_dldlt_batched": ([vp, c_int64, vp, vp, C.c_int32, vp, vp, vp, vp, vp], c_int),
    "eamd_dldlt_solve_batched": ([vp, c_int64, vp, vp, C.c_int32, vp, vp, vp, vp, vp], c_int),
    "eamd_spotrf_solve_batched_noalloc": ([vp, c_int64, vp, vp, vp, vp, vp, vp], c_int),
    "eamd_batched_plan_create": ([vp, c_int64, vp, C.c_int32, C.POINTER(vp)], c_int),
    "eamd_batched_plan_destroy": ([vp], c_int),
    "eamd_batched_plan_lanes": ([vp, C.c_int32], c_int),
    "eamd_spotrf_solve_batched_planned": ([vp, vp, vp, vp, vp, vp, vp], c_int),
    "eamd_dgesv_ir_batched": ([vp, c_int64, vp, vp, vp, vp, c_int64, vp, vp, vp, vp], c_int),
    "eamd_batched_lu_plan_create": ([vp, c_int64, vp, c_int64, C.POINTER(vp)], c_int),
    "eamd_dgesv_ir_batched_planned": ([vp, vp, vp, vp, vp, vp, vp, vp, vp, vp], c_int),
    "eamd_host_dgetrf": ([vp, c_int64, c_int64, vp, c_int64, vp, C.POINTER(C.c_int32),
                          C.POINTER(c_int64)], c_int),
    "eamd_host_dgetrs": ([vp, c_char, c_int64, c_int64, vp, c_int64, vp, vp, c_int64], c_int),
    "eamd_host_dpotrf": ([vp, c_char, c_int64, vp, c_int64, C.POINTER(c_int64)], c_int),
    "eamd_host_dpotrs": ([vp, c_char, c_int64, c_int64, vp, c_int64, vp, c_int64], c_int),
    "eamd_host_sgetrf": ([vp, c_int64, c_int64, vp, c_int64, vp, C.POINTER(C.c_int32),
                          C.POINTER(c_int64)], c_int),
    "eamd_host_sgetrs": ([vp, c_char, c_int64, c_int64, vp, c_int64, vp, vp, c_int64], c_int),
    "eamd_host_spotrf": ([vp, c_char, c_int64, vp, c_int64, C.POINTER(c_int64)], c_int),
    "eamd_host_spotrs": ([vp, c_char, c_int64, c_int64, vp, c_int64, vp, c_int64], c_int),
    "eamd_host_dgemm": ([vp, c_char, c_char, c_int64, c_int64, c_int64, c_dbl, vp, c_int64, vp,
                         c_int64, c_dbl, vp, c_int64], c_int),
    "eamd_host_dsyrk": ([vp, c_char, c_char, c_int64, c_int64, c_dbl, vp, c_int64, c_dbl, vp,
                         c_int64], c_int),
    "eamd_host_dtrsm": ([vp, c_char, c_char, c_char, c_char, c_int64, c_int64, c_dbl, vp,
                         c_int64, vp, c_int64], c_int),
    "eamd_host_dgesv_ir": ([vp, c_int64, c_int64, vp, c_int64, vp, c_int64, vp, c_int64,
                            C.POINTER(IrOpts), C.POINTER(IrReport)], c_int),
    "eamd_malloc": ([C.POINTER(vp), c_int64], c_int),
    "eamd_free": ([vp], c_int),
    "eamd_memcpy_htod": ([vp, vp, vp, c_int64], c_int),
    "eamd_memcpy_dtoh": ([vp, vp, vp, c_int64], c_int),
}


class EamdError(RuntimeError):
    """A non-zero status from the C ABI."""

    def __init__(self, fn, status, msg=""):
        self.status = status
        super().__init__(f"{fn} returned {status}" + (f": {msg}" if msg else ""))


def declared_symbols():
    """Names of the functions declared in include/eigen_amd.h."""
    with open(HEADER) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(eamd_\w+)\s*\(", text)))


def lib():
    """The loaded library; raises if it was not built (no fallback)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        f"{LIB_PATH} is missing: build it with `make -C eigen_amd/csrc` "
                        "(there is no CPU fallback)")
                L = C.CDLL(LIB_PATH)
                for name, (args, res) in _SIGS.items():
                    f = getattr(L, name)
                    f.argtypes = args
                    f.restype = res
                _lib = L
    return _lib


def check(fn, status):
    if status != 0:
        msg = lib().eamd_last_error()
        raise EamdError(fn, status, msg.decode() if msg else "")
    return status
