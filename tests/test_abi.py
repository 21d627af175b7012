"""CPU tests of the C ABI boundary: the library loads, exports every symbol
include/eigen_amd.h declares, and rejects bad arguments with the
reference's xerbla numbering before touching the GPU (lapack/lu.cpp:16-24,
lapack/cholesky.cpp:17-24, blas/level3_impl.h:47-57)."""
import ctypes as C
import os
import re

import pytest

import eigen_amd
from eigen_amd import _lib


def test_library_loads_and_exports_every_declared_symbol():
    L = eigen_amd.lib()
    names = _lib.declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes signatures cover the header
    assert set(names) <= set(_lib._SIGS)


def test_version():
    assert eigen_amd.lib().eamd_version() == 10000


NULL = None
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dgetrf_argument_codes():
    L = eigen_amd.lib()
    nbt = C.c_int32()
    fz = C.c_int64()
    assert L.eamd_dgetrf(NULL, -1, 4, NULL, 4, NULL, C.byref(nbt), C.byref(fz)) == -1
    assert L.eamd_dgetrf(NULL, 4, -1, NULL, 4, NULL, C.byref(nbt), C.byref(fz)) == -2
    assert L.eamd_dgetrf(NULL, 4, 4, NULL, 3, NULL, C.byref(nbt), C.byref(fz)) == -4


def test_dgetrs_argument_codes():
    L = eigen_amd.lib()
    assert L.eamd_dgetrs(NULL, b"X", 4, 1, NULL, 4, NULL, NULL, 4) == -1
    assert L.eamd_dgetrs(NULL, b"N", -1, 1, NULL, 4, NULL, NULL, 4) == -2
    assert L.eamd_dgetrs(NULL, b"N", 4, -1, NULL, 4, NULL, NULL, 4) == -3
    assert L.eamd_dgetrs(NULL, b"N", 4, 1, NULL, 3, NULL, NULL, 4) == -5
    assert L.eamd_dgetrs(NULL, b"N", 4, 1, NULL, 4, NULL, NULL, 3) == -8


def test_dpotrf_dpotrs_argument_codes():
    L = eigen_amd.lib()
    fc = C.c_int64()
    assert L.eamd_dpotrf(NULL, b"X", 4, NULL, 4, C.byref(fc)) == -1
    assert L.eamd_dpotrf(NULL, b"L", -1, NULL, 4, C.byref(fc)) == -2
    assert L.eamd_dpotrf(NULL, b"L", 4, NULL, 3, C.byref(fc)) == -4
    assert L.eamd_dpotrs(NULL, b"U", 4, -2, NULL, 4, NULL, 4) == -3
    assert L.eamd_dpotrs(NULL, b"U", 4, 1, NULL, 4, NULL, 2) == -7


def test_dgemm_dtrsm_dsyrk_argument_codes():
    L = eigen_amd.lib()
    assert L.eamd_dgemm(NULL, b"X", b"N", 2, 2, 2, 1.0, NULL, 2, NULL, 2, 0.0, NULL, 2) == -1
    assert L.eamd_dgemm(NULL, b"N", b"Q", 2, 2, 2, 1.0, NULL, 2, NULL, 2, 0.0, NULL, 2) == -2
    assert L.eamd_dgemm(NULL, b"N", b"N", 4, 2, 3, 1.0, NULL, 3, NULL, 3, 0.0, NULL, 4) == -8
    assert L.eamd_dgemm(NULL, b"N", b"N", 4, 2, 3, 1.0, NULL, 4, NULL, 2, 0.0, NULL, 4) == -10
    assert L.eamd_dgemm(NULL, b"N", b"N", 4, 2, 3, 1.0, NULL, 4, NULL, 3, 0.0, NULL, 3) == -13
    assert L.eamd_dtrsm(NULL, b"X", b"L", b"N", b"N", 2, 2, 1.0, NULL, 2, NULL, 2) == -1
    assert L.eamd_dtrsm(NULL, b"L", b"X", b"N", b"N", 2, 2, 1.0, NULL, 2, NULL, 2) == -2
    assert L.eamd_dtrsm(NULL, b"R", b"L", b"N", b"N", 5, 3, 1.0, NULL, 2, NULL, 5) == -9
    assert L.eamd_dsyrk(NULL, b"L", b"N", 5, 3, 1.0, NULL, 4, 0.0, NULL, 5) == -7


def test_null_handle_after_valid_arguments_is_an_error():
    L = eigen_amd.lib()
    fc = C.c_int64()
    assert L.eamd_dpotrf(NULL, b"L", 4, NULL, 4, C.byref(fc)) == -1
    assert b"handle" in L.eamd_last_error()


def test_ir_defaults():
    o = _lib.IrOpts()
    eigen_amd.lib().eamd_ir_opts_default(C.byref(o))
    assert o.tol == pytest.approx(1e-14)
    assert o.factor_precision == 4 and o.gmres == 1


# ------------------------------------------------ Fortran-ABI layer (CPU) --
LAPACK_H = os.path.join(ROOT, "include", "eigen_amd_lapack.h")
LAPACK_SO = os.path.join(ROOT, "eigen_amd", "libeigen_amd_lapack.so")


def _lapack():
    return C.CDLL(LAPACK_SO)


def test_lapack_layer_exports_every_declared_symbol():
    text = re.sub(r"/\*.*?\*/", "", open(LAPACK_H).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(d\w+_)\s*\(", text)))
    assert names == ["dgemm_", "dgetrf_", "dgetrs_", "dpotrf_", "dpotrs_", "dsyrk_", "dtrsm_"]
    L = _lapack()
    assert all(hasattr(L, n) for n in names)


def _ip(v):
    return C.byref(C.c_int(v))


def test_lapack_layer_argument_codes():
    """info = -(index of the illegal argument), checked before any device
    work (lapack/lu.cpp:16-24, lapack/cholesky.cpp:17-24)."""
    L = _lapack()
    info = C.c_int(0)
    piv = (C.c_int * 4)()
    L.dgetrf_(_ip(-1), _ip(4), None, _ip(4), piv, C.byref(info))
    assert info.value == -1
    L.dgetrf_(_ip(4), _ip(-1), None, _ip(4), piv, C.byref(info))
    assert info.value == -2
    L.dgetrf_(_ip(4), _ip(4), None, _ip(3), piv, C.byref(info))
    assert info.value == -4
    L.dgetrs_(C.c_char_p(b"X"), _ip(4), _ip(1), None, _ip(4), piv, None, _ip(4), C.byref(info))
    assert info.value == -1
    L.dgetrs_(C.c_char_p(b"N"), _ip(4), _ip(1), None, _ip(4), piv, None, _ip(3), C.byref(info))
    assert info.value == -8
    L.dpotrf_(C.c_char_p(b"Q"), _ip(4), None, _ip(4), C.byref(info))
    assert info.value == -1
    L.dpotrf_(C.c_char_p(b"L"), _ip(4), None, _ip(2), C.byref(info))
    assert info.value == -4
    L.dpotrs_(C.c_char_p(b"L"), _ip(4), _ip(-1), None, _ip(4), None, _ip(4), C.byref(info))
    assert info.value == -3
    # dsyrk_ reports through xerbla only (blas/level3_impl.h:381-389)
    one = C.byref(C.c_double(1.0))
    L.dsyrk_(C.c_char_p(b"X"), C.c_char_p(b"N"), _ip(4), _ip(2), one, None, _ip(4), one, None,
             _ip(4))
    # empty problems return at once with info 0
    L.dgetrf_(_ip(0), _ip(0), None, _ip(1), piv, C.byref(info))
    assert info.value == 0


def test_ldlt_argument_codes():
    L = eigen_amd.lib()
    assert L.eamd_dldlt_batched(NULL, -1, NULL, NULL, 8, NULL, NULL, NULL, NULL, NULL) == -2
    assert L.eamd_dldlt_batched(NULL, 4, NULL, NULL, 65, NULL, NULL, NULL, NULL, NULL) == -5
    assert L.eamd_dldlt_batched(NULL, 4, NULL, NULL, -1, NULL, NULL, NULL, NULL, NULL) == -5
    assert L.eamd_dldlt_solve_batched(NULL, -3, NULL, NULL, 8, NULL, NULL, NULL, NULL, NULL) == -2
    assert L.eamd_dldlt_solve_batched(NULL, 3, NULL, NULL, 70, NULL, NULL, NULL, NULL, NULL) == -5
