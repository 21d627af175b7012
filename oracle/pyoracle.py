"""ctypes bindings for the CPU checker (test infrastructure only).

Loads ``oracle/liboracle.so`` (the C restatement of Eigen's algorithms,
``oracle.c``) and, when it was built, ``oracle/_ref/libeigen_ref.so`` (the
reference compiled from its own headers, ``ref_eigen.cpp``).  Only tests,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
module; the product package never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ORC = None
_REF = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_d1 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f1 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
L = C.c_long


def build():
    """Compile the restatement (and the reference shim where headers exist)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def orc():
    global _ORC
    if _ORC is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        lib.orc_lu_blocked_d.argtypes = [L, L, _dp, L, _ip, C.POINTER(C.c_int), L, C.c_void_p]
        lib.orc_lu_blocked_d.restype = L
        lib.orc_lu_blocked_f.argtypes = [L, L, _fp, L, _ip, C.POINTER(C.c_int), L]
        lib.orc_lu_blocked_f.restype = L
        lib.orc_lu_unblocked_d.argtypes = [L, L, _dp, L, _ip, C.POINTER(C.c_int), C.c_void_p]
        lib.orc_lu_unblocked_d.restype = L
        lib.orc_transpositions_to_perm.argtypes = [L, _ip, _ip]
        lib.orc_lu_solve_d.argtypes = [L, L, _dp, L, _ip, _dp, L]
        lib.orc_lu_solve_f.argtypes = [L, L, _fp, L, _ip, _fp, L]
        for s, p in (("d", _dp), ("f", _fp)):
            getattr(lib, "orc_llt_unblocked_" + s).argtypes = [L, p, L]
            getattr(lib, "orc_llt_unblocked_" + s).restype = L
            getattr(lib, "orc_llt_blocked_" + s).argtypes = [L, p, L]
            getattr(lib, "orc_llt_blocked_" + s).restype = L
            getattr(lib, "orc_llt_solve_" + s).argtypes = [L, L, p, L, p, L]
        lib.orc_gemm_d.argtypes = [C.c_char, C.c_char, L, L, L, C.c_double, _dp, L, _dp, L,
                                   C.c_double, _dp, L]
        lib.orc_trsm_d.argtypes = [C.c_char, C.c_char, C.c_char, C.c_char, L, L, C.c_double,
                                   _dp, L, _dp, L]
        lib.orc_syrk_d.argtypes = [C.c_char, C.c_char, L, L, C.c_double, _dp, L, C.c_double,
                                   _dp, L]
        lib.orc_backward_error_d.argtypes = [L, L, _dp, L, _dp, L, _dp, L]
        lib.orc_backward_error_d.restype = C.c_double
        lib.orc_batched_llt_solve_f.argtypes = [L, _i64p, _ip, _f1, _i64p, _f1, _ip]
        lib.orc_llt_rank_update_d.argtypes = [L, _dp, L, _dp, C.c_double]
        lib.orc_llt_rank_update_d.restype = L
        lib.orc_ldlt_d.argtypes = [L, _dp, L, _ip, C.POINTER(C.c_int)]
        lib.orc_ldlt_d.restype = L
        lib.orc_ldlt_solve_d.argtypes = [L, L, _dp, L, _ip, _dp, L]
        lib.orc_ldlt_rank_update_d.argtypes = [L, _dp, L, _ip, _dp, C.c_double]
        lib.orc_ldlt_rank_update_d.restype = None
        lib.orc_batched_lu_solve_d.argtypes = [L, _i64p, _ip, _d1, _i64p, L, _d1, _ip]
        _ORC = lib
    return _ORC


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libeigen_ref.so"))


def ref():
    """The compiled reference (Eigen itself); None when it was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libeigen_ref.so")
        if not os.path.exists(path):
            return None
        lib = C.CDLL(path)
        lib.eref_version.restype = C.c_int
        lib.eref_set_threads.argtypes = [C.c_int]
        lib.eref_max_threads.restype = C.c_int
        lib.eref_lu_blocked_d.argtypes = [L, L, _dp, L, _ip, C.POINTER(C.c_int)]
        lib.eref_lu_blocked_d.restype = L
        lib.eref_lu_blocked_f.argtypes = [L, L, _fp, L, _ip, C.POINTER(C.c_int)]
        lib.eref_lu_blocked_f.restype = L
        lib.eref_partialpivlu_d.argtypes = [L, _dp, L, L, _dp, L, _dp, _ip, _dp, L,
                                            C.POINTER(C.c_double)]
        lib.eref_llt_d.argtypes = [L, _dp, L, L, _dp, L, _dp, _dp, L]
        lib.eref_llt_f.argtypes = [L, _fp, L, L, _fp, L, _fp, _fp, L]
        lib.eref_lu_rcond_d.argtypes = [L, _dp, L]
        lib.eref_lu_rcond_d.restype = C.c_double
        lib.eref_lu_inverse_d.argtypes = [L, _dp, L, _dp]
        lib.eref_llt_rcond_d.argtypes = [L, _dp, L]
        lib.eref_llt_rcond_d.restype = C.c_double
        lib.eref_llt_blocked_d.argtypes = [L, _dp, L]
        lib.eref_llt_blocked_d.restype = L
        lib.eref_llt_rank_update_d.argtypes = [L, _dp, L, _dp, C.c_double]
        lib.eref_llt_rank_update_d.restype = L
        lib.eref_ldlt_d.argtypes = [L, _dp, L, _ip, C.POINTER(C.c_int)]
        lib.eref_ldlt_d.restype = L
        lib.eref_ldlt_solve_d.argtypes = [L, L, _dp, L, _dp, L]
        lib.eref_ldlt_rank_update_d.argtypes = [L, _dp, L, _dp, C.c_double, _ip]
        lib.eref_ldlt_rank_update_d.restype = None
        lib.eref_gemm_d.argtypes = [L, L, L, _dp, L, _dp, L, _dp, L]
        lib.eref_trsm_d.argtypes = [C.c_char, C.c_char, C.c_char, C.c_char, L, L, _dp, L, _dp, L]
        lib.eref_syrk_lower_d.argtypes = [L, L, C.c_double, _dp, L, _dp, L]
        lib.eref_batched_llt_solve_f.argtypes = [L, _i64p, _ip, _f1, _i64p, _f1, _ip]
        lib.eref_batched_lu_solve_d.argtypes = [L, _i64p, _ip, _d1, _i64p, L, _d1, _ip]
        _REF = lib
    return _REF


# ---------------------------------------------------------------- helpers --
def lu(a, max_block=256, margins=False):
    """Restated blocked_lu on a copy of ``a``; returns (lu, transp, nbt, fz[, margin])."""
    a = np.array(a, dtype=np.float64, order="F", copy=True)
    m, n = a.shape
    t = np.zeros(min(m, n), dtype=np.int32)
    nbt = C.c_int(0)
    mg = np.zeros(min(m, n), dtype=np.float64) if margins else None
    fz = orc().orc_lu_blocked_d(m, n, a, max(m, 1), t, C.byref(nbt), max_block,
                                mg.ctypes.data if margins else None)
    out = (a, t, nbt.value, int(fz))
    return out + (mg,) if margins else out


def lu_f(a, max_block=256):
    a = np.array(a, dtype=np.float32, order="F", copy=True)
    m, n = a.shape
    t = np.zeros(min(m, n), dtype=np.int32)
    nbt = C.c_int(0)
    fz = orc().orc_lu_blocked_f(m, n, a, max(m, 1), t, C.byref(nbt), max_block)
    return a, t, nbt.value, int(fz)


def perm_from_transpositions(t):
    p = np.zeros(len(t), dtype=np.int32)
    orc().orc_transpositions_to_perm(len(t), np.ascontiguousarray(t, dtype=np.int32), p)
    return p


def lu_solve(lu_, t, b):
    b = np.array(b, dtype=np.float64, order="F", copy=True)
    if b.ndim == 1:
        b = b.reshape(-1, 1, order="F")
    n = lu_.shape[0]
    orc().orc_lu_solve_d(n, b.shape[1], np.asfortranarray(lu_), n,
                         np.ascontiguousarray(t, dtype=np.int32), b, n)
    return b


def llt_rank_update(lfac, v, sigma, reference=False):
    """LLT rank-one update of a lower factor (LLT.h:231-299) on a copy;
    returns (factor, info) with info -1 or the failing column.  reference=True
    runs Eigen itself (oracle/_ref)."""
    f = np.asfortranarray(np.array(lfac, dtype=np.float64))
    vv = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
    n = f.shape[0]
    lib = ref() if reference else orc()
    r = (lib.eref_llt_rank_update_d if reference else lib.orc_llt_rank_update_d)(
        n, f, n, vv, float(sigma))
    return f, int(r)


def ldlt(a, reference=False):
    """LDLT<MatrixXd, Lower> (LDLT.h:290-395, 491-522) of a copy of a:
    returns (matrixLDLT with the strict upper part of a kept, transpositions,
    sign 0..3 = PositiveSemiDef, NegativeSemiDef, ZeroSign, Indefinite,
    ok).  reference=True runs Eigen itself (oracle/_ref)."""
    f = np.array(a, dtype=np.float64, order="F", copy=True)
    n = f.shape[0]
    t = np.zeros(max(n, 1), dtype=np.int32)
    sg = C.c_int(-1)
    lib = ref() if reference else orc()
    r = (lib.eref_ldlt_d if reference else lib.orc_ldlt_d)(n, f, max(n, 1), t, C.byref(sg))
    return f, t[:n], int(sg.value), bool(r)


def ldlt_solve(f, t, b):
    """x = A^{-1} b from ldlt()'s factors (LDLT.h:559-593)."""
    b = np.array(b, dtype=np.float64, order="F", copy=True)
    vec = b.ndim == 1
    if vec:
        b = b.reshape(-1, 1, order="F")
    n = f.shape[0]
    orc().orc_ldlt_solve_d(n, b.shape[1], np.asfortranarray(f), max(n, 1),
                           np.ascontiguousarray(t, dtype=np.int32), b, max(n, 1))
    return b.reshape(-1) if vec else b


def ldlt_solve_reference(a, b):
    """Eigen's LDLT<MatrixXd, Lower>(a).solve(b), run by oracle/_ref."""
    b = np.array(b, dtype=np.float64, order="F", copy=True)
    vec = b.ndim == 1
    if vec:
        b = b.reshape(-1, 1, order="F")
    a = np.asfortranarray(np.array(a, dtype=np.float64))
    n = a.shape[0]
    ref().eref_ldlt_solve_d(n, b.shape[1], a, max(n, 1), b, max(n, 1))
    return b.reshape(-1) if vec else b


def ldlt_rank_update(f, t, w, sigma):
    """LDLT rank-one update (LDLT.h:403-449) of ldlt()'s factor f with
    transpositions t, on a copy: the factor of P^T L D L^T P + sigma w w^T."""
    f = np.array(f, dtype=np.float64, order="F", copy=True)
    n = f.shape[0]
    orc().orc_ldlt_rank_update_d(n, f, max(n, 1), np.ascontiguousarray(t, dtype=np.int32),
                                 np.ascontiguousarray(w, dtype=np.float64), float(sigma))
    return f


def ldlt_rank_update_reference(a, w, sigma):
    """Eigen's LDLT<MatrixXd, Lower>(a).rankUpdate(w, sigma), run by
    oracle/_ref: (matrixLDLT with a's strict upper part kept, transpositions)."""
    f = np.array(a, dtype=np.float64, order="F", copy=True)
    n = f.shape[0]
    t = np.zeros(max(n, 1), dtype=np.int32)
    ref().eref_ldlt_rank_update_d(n, f, max(n, 1), np.ascontiguousarray(w, dtype=np.float64),
                                  float(sigma), t)
    return f, t[:n]


def llt(a, dtype=np.float64):
    a = np.array(a, dtype=dtype, order="F", copy=True)
    n = a.shape[0]
    f = orc().orc_llt_blocked_d if dtype == np.float64 else orc().orc_llt_blocked_f
    r = f(n, a, max(n, 1))
    return a, int(r)


def llt_solve(l, b):
    dtype = l.dtype
    b = np.array(b, dtype=dtype, order="F", copy=True)
    if b.ndim == 1:
        b = b.reshape(-1, 1, order="F")
    n = l.shape[0]
    f = orc().orc_llt_solve_d if dtype == np.float64 else orc().orc_llt_solve_f
    f(n, b.shape[1], np.asfortranarray(l), n, b, n)
    return b


def gemm(a, b, c=None, alpha=1.0, beta=0.0, transa="N", transb="N"):
    a = np.asfortranarray(a, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64)
    m = a.shape[0] if transa == "N" else a.shape[1]
    k = a.shape[1] if transa == "N" else a.shape[0]
    n = b.shape[1] if transb == "N" else b.shape[0]
    c = np.zeros((m, n), order="F") if c is None else np.array(c, order="F", dtype=np.float64)
    orc().orc_gemm_d(transa.encode(), transb.encode(), m, n, k, alpha, a, max(a.shape[0], 1),
                     b, max(b.shape[0], 1), beta, c, max(m, 1))
    return c


def trsm(side, uplo, trans, diag, a, b, alpha=1.0):
    b = np.array(b, dtype=np.float64, order="F", copy=True)
    a = np.asfortranarray(a, dtype=np.float64)
    m, n = b.shape
    orc().orc_trsm_d(side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n, alpha,
                     a, max(a.shape[0], 1), b, max(m, 1))
    return b


def syrk(uplo, trans, a, c, alpha=1.0, beta=1.0):
    a = np.asfortranarray(a, dtype=np.float64)
    c = np.array(c, dtype=np.float64, order="F", copy=True)
    n = c.shape[0]
    k = a.shape[1] if trans == "N" else a.shape[0]
    orc().orc_syrk_d(uplo.encode(), trans.encode(), n, k, alpha, a, max(a.shape[0], 1), beta,
                     c, max(n, 1))
    return c


def backward_error(a, x, b):
    a = np.asfortranarray(a, dtype=np.float64)
    x = np.asfortranarray(x.reshape(a.shape[0], -1), dtype=np.float64)
    b = np.asfortranarray(b.reshape(a.shape[0], -1), dtype=np.float64)
    n = a.shape[0]
    return orc().orc_backward_error_d(n, x.shape[1], a, n, x, n, b, n)


def batched_llt_solve_f(offA, n, A, offb, b):
    b = np.array(b, dtype=np.float32, copy=True)
    info = np.zeros(len(n), dtype=np.int32)
    orc().orc_batched_llt_solve_f(len(n), offA, n, A, offb, b, info)
    return b, info


def batched_lu_solve_d(offA, n, A, offb, nrhs, B):
    B = np.array(B, dtype=np.float64, copy=True)
    info = np.zeros(len(n), dtype=np.int32)
    orc().orc_batched_lu_solve_d(len(n), offA, n, A, offb, nrhs, B, info)
    return B, info


# ------------------------------------------------------- condition numbers --
def inv_l1_norm_estimate(n, solve, solve_adj):
    """Estimate of ||A^-1||_1 from solves with A and A^T (Hager's method with
    Higham's refinements, N. J. Higham, "FORTRAN codes for estimating the
    one-norm of a real or complex matrix", ACM TOMS 14, 1988, Algorithms 4.1
    and 5.1; restated from rcond_invmatrix_L1_norm_estimate,
    Eigen/src/Core/ConditionEstimator.h:56-141).  `solve(v)` returns A^-1 v,
    `solve_adj(v)` returns A^-T v (1-D float64 arrays)."""
    if n == 0:
        return 0.0
    v = solve(np.full(n, 1.0 / n))
    lower = float(np.abs(v).sum())
    if n == 1:
        return lower
    old_lower, old_sign, old_idx = lower, None, -1
    for k in range(4):
        sign = np.where(v < 0, -1.0, 1.0)  # rcond_compute_sign, real case (:21-27)
        if k > 0 and np.array_equal(sign, old_sign):
            break
        z = solve_adj(sign)
        idx = int(np.argmax(np.abs(z)))  # first maximum, as maxCoeff
        if idx == old_idx:
            break
        e = np.zeros(n)
        e[idx] = 1.0
        v = solve(e)
        lower = float(np.abs(v).sum())
        if lower <= old_lower:
            break
        old_sign, old_idx, old_lower = sign, idx, lower
    # Higham's alternating-sign vector (:126-137)
    i = np.arange(n, dtype=np.float64)
    alt = np.where(np.arange(n) % 2 == 0, 1.0, -1.0) * (1.0 + i / (n - 1))
    alt_lower = 2.0 * float(np.abs(solve(alt)).sum()) / (3.0 * n)
    return max(lower, alt_lower)


def rcond_from(norm1, n, solve, solve_adj):
    """1 / (||A||_1 ||A^-1||_1 estimate) (ConditionEstimator.h:157-172)."""
    if n == 0:
        return 1.0
    if norm1 == 0.0:
        return 0.0
    if n == 1:
        return 1.0
    inv = inv_l1_norm_estimate(n, solve, solve_adj)
    return 0.0 if inv == 0.0 else (1.0 / inv) / norm1


def lu_rcond(a):
    """PartialPivLU::rcond on the oracle's factorization (PartialPivLU.h:184-188):
    ||A||_1 is the maximum absolute column sum of A (:522)."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    lu_, t, _, _ = lu(a)
    L = np.tril(lu_, -1) + np.eye(n)
    U = np.triu(lu_)
    perm = perm_from_transpositions(t)  # P A = L U, (P b)[perm[i]] = b[i]
    import scipy.linalg as sl

    def solve(b):
        pb = np.empty(n)
        pb[perm] = b
        return sl.solve_triangular(U, sl.solve_triangular(L, pb, lower=True, unit_diagonal=True))

    def solve_adj(b):
        y = sl.solve_triangular(L.T, sl.solve_triangular(U.T, b, lower=True), lower=False,
                                unit_diagonal=True)
        return y[perm]

    return rcond_from(float(np.abs(a).sum(axis=0).max()), n, solve, solve_adj)


def llt_rcond(a):
    """LLT::rcond (LLT.h:161-166): ||A||_1 of the selfadjoint matrix read from
    its lower triangle (LLT.h:434-445)."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    l_, fail = llt(a)
    if fail >= 0:
        return -1.0
    L = np.tril(l_)
    low = np.tril(a)
    sym = low + np.tril(a, -1).T
    import scipy.linalg as sl

    def solve(b):
        return sl.solve_triangular(L.T, sl.solve_triangular(L, b, lower=True), lower=False)

    return rcond_from(float(np.abs(sym).sum(axis=0).max()), n, solve, solve)

