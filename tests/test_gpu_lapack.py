"""GPU tests of the Fortran-ABI layer (libeigen_amd_lapack.so): dgetrf_,
dgetrs_, dpotrf_, dpotrs_, dgemm_, dtrsm_ on host column-major arrays,
against the oracle (1-based ipiv equal to the oracle's transpositions + 1,
info rules of lapack/lu.cpp:14-41 and lapack/cholesky.cpp:14-36)."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import inputs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def L():
    return C.CDLL(os.path.join(ROOT, "eigen_amd", "libeigen_amd_lapack.so"))


def _i(v):
    return C.byref(C.c_int(v))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def berr(a, x, b):
    r = b - a @ x
    return (np.abs(r).max(axis=0) /
            (np.abs(a).sum(axis=1).max() * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))).max()


@pytest.mark.parametrize("n", [1, 37, 300])
def test_dgetrf_dgetrs(gpu, L, n):
    a0 = inputs.normal(41, n, n)
    a = np.asfortranarray(a0.copy())
    ipiv = np.zeros(n, dtype=np.int32)
    info = C.c_int(-7)
    L.dgetrf_(_i(n), _i(n), _d(a), _i(n), ipiv.ctypes.data_as(C.POINTER(C.c_int)), C.byref(info))
    assert info.value == 0
    lu_r, t_r, _, _ = po.lu(a0)
    np.testing.assert_array_equal(ipiv, t_r + 1)
    np.testing.assert_allclose(a, lu_r, rtol=0, atol=1e-11 * np.abs(lu_r).max())
    b0 = np.asfortranarray(inputs.normal(42, n, 3))
    for trans, op in ((b"N", a0), (b"T", a0.T)):
        b = np.asfortranarray(b0.copy())
        L.dgetrs_(C.c_char_p(trans), _i(n), _i(3), _d(a), _i(n),
                  ipiv.ctypes.data_as(C.POINTER(C.c_int)), _d(b), _i(n), C.byref(info))
        assert info.value == 0
        assert berr(op, b, b0) <= 1e-14


def test_dgetrf_zero_pivot_info(gpu, L):
    n = 20
    a = np.asfortranarray(inputs.normal(43, n, n))
    a[:, 6] = 0.0
    ipiv = np.zeros(n, dtype=np.int32)
    info = C.c_int(0)
    L.dgetrf_(_i(n), _i(n), _d(a), _i(n), ipiv.ctypes.data_as(C.POINTER(C.c_int)), C.byref(info))
    assert info.value == 7  # first exact-zero pivot (0-based 6) + 1


def test_dpotrf_dpotrs(gpu, L):
    n = 200
    s0 = inputs.spd(44, n)
    for uplo in (b"L", b"U"):
        a = np.asfortranarray(s0.copy())
        info = C.c_int(-7)
        L.dpotrf_(C.c_char_p(uplo), _i(n), _d(a), _i(n), C.byref(info))
        assert info.value == 0
        f = np.tril(a) if uplo == b"L" else np.triu(a).T
        np.testing.assert_allclose(f @ f.T, s0, rtol=0, atol=1e-12 * np.abs(s0).max())
        b0 = np.asfortranarray(inputs.normal(45, n, 2))
        b = np.asfortranarray(b0.copy())
        L.dpotrs_(C.c_char_p(uplo), _i(n), _i(2), _d(a), _i(n), _d(b), _i(n), C.byref(info))
        assert info.value == 0
        assert berr(s0, b, b0) <= 1e-14
    bad = np.asfortranarray(s0.copy())
    bad[10, 10] = -1.0
    L.dpotrf_(C.c_char_p(b"L"), _i(n), _d(bad), _i(n), C.byref(info))
    assert info.value > 0


def test_dgemm_dtrsm(gpu, L):
    m, n, k = 70, 50, 30
    a = np.asfortranarray(inputs.normal(46, m, k))
    b = np.asfortranarray(inputs.normal(47, k, n))
    c0 = np.asfortranarray(inputs.normal(48, m, n))
    c = np.asfortranarray(c0.copy())
    alpha, beta = C.c_double(0.5), C.c_double(-2.0)
    L.dgemm_(C.c_char_p(b"N"), C.c_char_p(b"N"), _i(m), _i(n), _i(k), C.byref(alpha), _d(a),
             _i(m), _d(b), _i(k), C.byref(beta), _d(c), _i(m))
    np.testing.assert_allclose(c, 0.5 * a @ b - 2.0 * c0, rtol=0, atol=1e-12 * 50)
    t = np.asfortranarray(inputs.normal(49, m, m))
    t[np.diag_indices(m)] += 10.0
    x0 = np.asfortranarray(inputs.normal(50, m, n))
    x = np.asfortranarray(x0.copy())
    one = C.c_double(1.0)
    L.dtrsm_(C.c_char_p(b"L"), C.c_char_p(b"L"), C.c_char_p(b"N"), C.c_char_p(b"N"), _i(m), _i(n),
             C.byref(one), _d(t), _i(m), _d(x), _i(m))
    np.testing.assert_allclose(np.tril(t) @ x, x0, rtol=0, atol=1e-11)
