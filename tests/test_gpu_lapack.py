"""GPU tests of the Fortran-ABI layer (libeigen_amd_lapack.so): dgetrf_,
dgetrs_, dpotrf_, dpotrs_ (and their FP32 s* forms), dgemm_, dtrsm_, dsyrk_
on host column-major arrays,
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


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def test_sgetrf_sgetrs_spotrf_spotrs(gpu, L):
    """The FP32 stubs (FP32 arrays, FP64 arithmetic inside): ipiv equal to
    the oracle's FP64 transpositions on the FP32 values + 1, each factor
    entry that factor rounded once, FP32-level backward errors, and the
    info rules of the FP64 stubs."""
    n = 150
    a0 = inputs.normal(141, n, n).astype(np.float32)
    a = np.asfortranarray(a0.copy())
    ipiv = np.zeros(n, dtype=np.int32)
    info = C.c_int(-7)
    L.sgetrf_(_i(n), _i(n), _f(a), _i(n), ipiv.ctypes.data_as(C.POINTER(C.c_int)), C.byref(info))
    assert info.value == 0
    A = a0.astype(np.float64)
    lu_r, t_r, _, _ = po.lu(A)
    np.testing.assert_array_equal(ipiv, t_r + 1)
    np.testing.assert_allclose(a.astype(np.float64), lu_r, rtol=2.0 ** -24,
                               atol=1e-11 * np.abs(lu_r).max())
    b0 = np.asfortranarray(inputs.normal(142, n, 3).astype(np.float32))
    for trans, op in ((b"N", A), (b"T", A.T)):
        b = np.asfortranarray(b0.copy())
        L.sgetrs_(C.c_char_p(trans), _i(n), _i(3), _f(a), _i(n),
                  ipiv.ctypes.data_as(C.POINTER(C.c_int)), _f(b), _i(n), C.byref(info))
        assert info.value == 0
        assert berr(op, b.astype(np.float64), b0.astype(np.float64)) <= 1e-5
    z = np.asfortranarray(a0.copy())
    z[:, 6] = 0.0
    L.sgetrf_(_i(n), _i(n), _f(z), _i(n), ipiv.ctypes.data_as(C.POINTER(C.c_int)), C.byref(info))
    assert info.value == 7
    s0 = inputs.spd(143, n).astype(np.float32)
    for uplo in (b"L", b"U"):
        s = np.asfortranarray(s0.copy())
        L.spotrf_(C.c_char_p(uplo), _i(n), _f(s), _i(n), C.byref(info))
        assert info.value == 0
        f = np.tril(s) if uplo == b"L" else np.triu(s).T
        f = f.astype(np.float64)
        S = s0.astype(np.float64)
        np.testing.assert_allclose(f @ f.T, S, rtol=0, atol=1e-6 * np.abs(S).max())
        b = np.asfortranarray(b0.copy())
        L.spotrs_(C.c_char_p(uplo), _i(n), _i(3), _f(s), _i(n), _f(b), _i(n), C.byref(info))
        assert info.value == 0
        assert berr(S, b.astype(np.float64), b0.astype(np.float64)) <= 1e-5
    bad = np.asfortranarray(s0.copy())
    bad[10, 10] = -1.0
    L.spotrf_(C.c_char_p(b"L"), _i(n), _f(bad), _i(n), C.byref(info))
    assert info.value == 11


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


@pytest.mark.parametrize("m,n", [(300, 200), (200, 300), (40, 13), (13, 40), (12, 16), (1, 5),
                                 (7, 1)])
def test_dgetrf_rectangular(gpu, L, m, n):
    """dgetrf_ on m x n (lapack/lu.cpp:14-41 -> blocked_lu(rows, cols)):
    min(m, n) pivots equal to Eigen's own blocked_lu (oracle/_ref) and the
    factors within rounding of its output, including the reference's
    treatment of a wide matrix's columns beyond m (left as they were above
    16 pivots, solved below)."""
    a0 = inputs.normal(51 + m, m, n)
    a = np.asfortranarray(a0.copy())
    k = min(m, n)
    ipiv = np.zeros(k, dtype=np.int32)
    info = C.c_int(-7)
    L.dgetrf_(_i(m), _i(n), _d(a), _i(m), ipiv.ctypes.data_as(C.POINTER(C.c_int)),
              C.byref(info))
    assert info.value == 0
    ref = np.asfortranarray(a0.copy())
    t_ref = np.zeros(k, dtype=np.int32)
    nbt = C.c_int(0)
    fz = po.ref().eref_lu_blocked_d(m, n, ref, m, t_ref, C.byref(nbt))
    assert fz == -1
    np.testing.assert_array_equal(ipiv, t_ref + 1)
    np.testing.assert_allclose(a, ref, rtol=0, atol=1e-11 * max(1.0, np.abs(ref).max()))
    if m < n and k > 16:
        np.testing.assert_array_equal(a[:, m:], a0[:, m:])
    else:
        # P A = L U over every column
        pa = a0.copy()
        for i in range(k):
            pa[[i, t_ref[i]]] = pa[[t_ref[i], i]]
        lo = np.tril(a[:, :k], -1) + np.eye(m, k)
        np.testing.assert_allclose(lo @ np.triu(a[:k, :]), pa, rtol=0, atol=1e-12 * n)


@pytest.mark.parametrize("uplo", [b"L", b"U"])
@pytest.mark.parametrize("op", [b"N", b"T"])
def test_dsyrk(gpu, L, uplo, op):
    """dsyrk_ (blas/level3_impl.h:357-391): the chosen triangle of
    alpha op(A) op(A)^T + beta C against the FP64 product, the other
    triangle untouched; k = 0 only scales by beta."""
    n, k = 90, 37
    a = np.asfortranarray(inputs.normal(60, n, k) if op == b"N" else inputs.normal(60, k, n))
    opa = a if op == b"N" else a.T
    c0 = np.asfortranarray(inputs.normal(61, n, n))
    c = np.asfortranarray(c0.copy())
    alpha, beta = C.c_double(0.75), C.c_double(-0.5)
    lda = n if op == b"N" else k
    L.dsyrk_(C.c_char_p(uplo), C.c_char_p(op), _i(n), _i(k), C.byref(alpha), _d(a), _i(lda),
             C.byref(beta), _d(c), _i(n))
    ref = 0.75 * opa @ opa.T - 0.5 * c0
    tri, other = (np.tril, lambda x: np.triu(x, 1)) if uplo == b"L" else \
        (np.triu, lambda x: np.tril(x, -1))
    np.testing.assert_allclose(tri(c), tri(ref), rtol=0, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_array_equal(other(c), other(c0))
    c = np.asfortranarray(c0.copy())
    L.dsyrk_(C.c_char_p(uplo), C.c_char_p(op), _i(n), _i(0), C.byref(alpha), _d(a), _i(lda),
             C.byref(beta), _d(c), _i(n))
    np.testing.assert_array_equal(tri(c), tri(-0.5 * c0))
    np.testing.assert_array_equal(other(c), other(c0))
