"""GPU test of the Eigen plug-point specialisations
(include/eigen_amd/eigen_plugin.h): unmodified Eigen user code
(tests/adapter/eigen_plugin_check.cpp, compiled with g++ against the Eigen
headers into oracle/_ref/) runs PartialPivLU<MatrixXd|MatrixXf>,
LLT<MatrixXd|MatrixXf>, GEMM and triangular solves through the C ABI; its outputs are checked
against the oracle and closed forms here."""
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "eigen_plugin_check")


def _load(d, name, shape, dtype=np.float64):
    a = np.fromfile(os.path.join(d, name), dtype=dtype)
    return a.reshape(shape, order="F") if len(shape) > 1 else a


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/eigen_plugin_check not built (needs the Eigen headers at build time)")
    d = str(tmp_path_factory.mktemp("plugin"))
    r = subprocess.run([BIN, d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    return d


def berr(a, x, b):
    r = b - a @ x
    return (np.abs(r).max(axis=0) /
            (np.abs(a).sum(axis=1).max() * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))).max()


def test_plugin_partialpivlu(gpu, out):
    n = 300
    A = _load(out, "lu_A.f64", (n, n))
    b = _load(out, "lu_b.f64", (n, 2))
    LU = _load(out, "lu_LU.f64", (n, n))
    perm = _load(out, "lu_perm.i32", (n,), np.int32)
    x = _load(out, "lu_x.f64", (n, 2))
    lu_r, t_r, nbt_r, fz_r = po.lu(A)
    np.testing.assert_array_equal(perm, po.perm_from_transpositions(t_r))
    np.testing.assert_allclose(LU, lu_r, rtol=0, atol=1e-11 * np.abs(lu_r).max())
    assert berr(A, x, b) <= 1e-14
    det = _load(out, "lu_det.f64", (1,))[0]
    det_r = (-1.0) ** nbt_r * np.prod(np.diag(lu_r))
    assert abs(det - det_r) <= 1e-9 * abs(det_r)


def test_plugin_llt(gpu, out):
    n = 300
    S = _load(out, "llt_S.f64", (n, n))
    b = _load(out, "llt_b.f64", (n, 3))
    L = _load(out, "llt_L.f64", (n, n))
    U = _load(out, "llt_U.f64", (n, n))
    x = _load(out, "llt_x.f64", (n, 3))
    info = _load(out, "llt_info.i32", (2,), np.int32)
    assert (info == 0).all()
    l_r, fail = po.llt(S)
    assert fail == -1
    l_r = np.tril(l_r)
    np.testing.assert_allclose(L, l_r, rtol=0, atol=1e-12 * np.abs(l_r).max())
    np.testing.assert_allclose(U, l_r.T, rtol=0, atol=1e-12 * np.abs(l_r).max())
    assert berr(S, x, b) <= 1e-14
    assert _load(out, "llt_bad_info.i32", (1,), np.int32)[0] == 1


def test_plugin_gemm(gpu, out):
    A = _load(out, "gemm_A.f64", (211, 173))
    B = _load(out, "gemm_B.f64", (173, 157))
    C0 = _load(out, "gemm_C0.f64", (211, 157))
    C = _load(out, "gemm_C.f64", (211, 157))
    ref = C0 + A @ B
    np.testing.assert_allclose(C, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    T = _load(out, "gemm_T.f64", (173, 64))
    assert T.shape == (173, 64) and np.isfinite(T).all()


def test_plugin_trsm(gpu, out):
    T = _load(out, "trsm_T.f64", (180, 180))
    X10 = _load(out, "trsm_X10.f64", (180, 40))
    X20 = _load(out, "trsm_X20.f64", (55, 180))
    X1 = _load(out, "trsm_X1.f64", (180, 40))
    X2 = _load(out, "trsm_X2.f64", (55, 180))
    np.testing.assert_allclose(np.tril(T) @ X1, X10, rtol=0, atol=1e-11 * np.abs(X10).max())
    Tu = _load(out, "trsm_Tu.f64", (180, 180))
    Uu = np.triu(Tu, 1) + np.eye(180)
    np.testing.assert_allclose(X2 @ Uu, X20, rtol=0, atol=1e-11 * np.abs(X20).max())


def _counts(d, name):
    return dict(zip(["getrf", "potrf", "gemm", "syrk", "trsm", "trsv"],
                    np.fromfile(os.path.join(d, name), dtype=np.int64)))


def test_plugin_vector_solves_reach_the_gpu(gpu, out):
    """lu.solve(VectorXd) and llt.solve(VectorXd) go through the vector
    triangular-solve plug point (TriangularSolverVector.h:88-139): two GPU
    solves each, backward error <= 1e-14."""
    n = 300
    A = _load(out, "lu_A.f64", (n, n))
    b = _load(out, "luv_b.f64", (n,))
    x = _load(out, "luv_x.f64", (n,))
    before, after = _counts(out, "cnt_lu.i64"), _counts(out, "cnt_luv.i64")
    assert after["trsv"] - before["trsv"] == 2
    assert po.backward_error(A, x, b) <= 1e-14
    S = _load(out, "llt_S.f64", (n, n))
    bs = _load(out, "llt_b.f64", (n, 3))[:, 0]
    for name in ("lltv_x.f64", "lltv_xu.f64"):
        assert po.backward_error(S, _load(out, name, (n,)), bs) <= 1e-14
    c = _counts(out, "cnt_llt.i64")
    assert c["trsv"] - after["trsv"] == 4 and c["potrf"] == 2


def test_plugin_syrk(gpu, out):
    """selfadjointView<Lower>().rankUpdate(A, 0.5) and
    selfadjointView<Upper>().rankUpdate(B^T, -2) run as GPU SYRKs: the
    updated triangle against the FP64 product, the other one untouched."""
    Sa = _load(out, "syrk_A.f64", (150, 70))
    Sb = _load(out, "syrk_B.f64", (60, 150))
    C0 = _load(out, "syrk_C0.f64", (150, 150))
    Sl = _load(out, "syrk_L.f64", (150, 150))
    Su = _load(out, "syrk_U.f64", (150, 150))
    rl = C0 + 0.5 * Sa @ Sa.T
    ru = C0 - 2.0 * Sb.T @ Sb
    tol = 1e-13 * max(np.abs(rl).max(), np.abs(ru).max())
    np.testing.assert_allclose(np.tril(Sl), np.tril(rl), rtol=0, atol=tol)
    np.testing.assert_array_equal(np.triu(Sl, 1), np.triu(C0, 1))
    np.testing.assert_allclose(np.triu(Su), np.triu(ru), rtol=0, atol=tol)
    np.testing.assert_array_equal(np.tril(Su, -1), np.tril(C0, -1))
    before, after = _counts(out, "cnt_trsm.i64"), _counts(out, "cnt_syrk.i64")
    assert after["syrk"] - before["syrk"] == 2


def test_plugin_fp32_factorizations(gpu, out):
    """PartialPivLU<MatrixXf> and LLT<MatrixXf> (lower, upper) reach the GPU
    through the float plug points, which compute in FP64 on the FP32 values:
    the pivots equal the oracle's FP64 pivots on those values, and every
    factor entry is the oracle's FP64 factor rounded once (a factorization
    carried out in FP32 would be off by many FP32 ulps).  Eigen's FP32
    substitution then solves with them to FP32 backward error, and a
    non-SPD matrix reports NumericalIssue."""
    n = 257
    f32 = np.float32
    A = _load(out, "luf_A.f32", (n, n), f32).astype(np.float64)
    b = _load(out, "luf_b.f32", (n, 2), f32).astype(np.float64)
    LU = _load(out, "luf_LU.f32", (n, n), f32).astype(np.float64)
    perm = _load(out, "luf_perm.i32", (n,), np.int32)
    x = _load(out, "luf_x.f32", (n, 2), f32).astype(np.float64)
    lu_r, t_r, _, _ = po.lu(A)
    np.testing.assert_array_equal(perm, po.perm_from_transpositions(t_r))
    # one FP32 rounding (half an ulp, 2^-24 relative) of an FP64 factor that
    # is itself within 1e-11 of the oracle's
    np.testing.assert_allclose(LU, lu_r, rtol=2.0 ** -24, atol=1e-11 * np.abs(lu_r).max())
    assert berr(A, x, b) <= 1e-5
    S = _load(out, "lltf_S.f32", (n, n), f32).astype(np.float64)
    bs = _load(out, "lltf_b.f32", (n, 3), f32).astype(np.float64)
    L = _load(out, "lltf_L.f32", (n, n), f32).astype(np.float64)
    U = _load(out, "lltf_U.f32", (n, n), f32).astype(np.float64)
    xs = _load(out, "lltf_x.f32", (n, 3), f32).astype(np.float64)
    info = _load(out, "lltf_info.i32", (3,), np.int32)
    np.testing.assert_array_equal(info, [0, 0, 1])
    l_r, fail = po.llt(S)
    assert fail == -1
    l_r = np.tril(l_r)
    tol = 1e-12 * np.abs(l_r).max()
    np.testing.assert_allclose(L, l_r, rtol=2.0 ** -24, atol=tol)
    # Eigen's FP32 G G^T is not exactly symmetric: the Upper factorization
    # reads the upper triangle, so its reference is the one of that triangle
    Su = np.triu(S) + np.triu(S, 1).T
    u_r, fail = po.llt(Su)
    assert fail == -1
    u_r = np.tril(u_r).T
    np.testing.assert_allclose(U, u_r, rtol=2.0 ** -24, atol=tol)
    assert berr(S, xs, bs) <= 1e-5
    before, after = _counts(out, "cnt_syrk.i64"), _counts(out, "cnt_f32.i64")
    assert after["getrf"] - before["getrf"] == 1 and after["potrf"] - before["potrf"] == 3


def test_plugin_empty_matrices(gpu, out):
    """PartialPivLU and LLT of a 0 x 0 matrix succeed, as in the reference."""
    assert _load(out, "empty_ok.i32", (1,), np.int32)[0] == 1
