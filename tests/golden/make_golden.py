"""Writes the golden fixtures of tests/golden/ from the reference itself.

Runs in the build container only: it needs oracle/_ref/libeigen_ref.so,
which oracle/Makefile compiles from the Eigen 3.3.90 headers (g++ -O3
-mavx2 -mfma; the snapshot's AVX-512 path crashes).  The reference's own
tests hold no golden vectors for this path (they are property tests), so
these fixtures are Eigen's outputs on seeded inputs (tests/inputs.py):

  lu.npz    PartialPivLU (blocked_lu, PartialPivLU.h:426-496) for a sweep of
            sizes: transpositions, nb_transpositions, first zero pivot, LU
            factors (n <= 64), solution of A x = b, determinant; plus the
            n = 1024 FP64 instance of the plain drop-in configuration.
  llt.npz   LLT<MatrixXd> (LLT.h:330-360): L (n <= 33), solution, info;
            forced non-SPD cases with their failing column.
  gemm.npz  C = A*B on the edge shapes of bench/perf_monitoring/
            gemm_settings.txt reduced to <= 240 per side (full C for the
            small ones; row sums, column sums and 256 sampled entries else).
  batched.npz  a 1,000-system sample of the batched LLT (FP32) and batched LU
            (FP64, 4 RHS, 48 systems) workloads with Eigen's per-system solutions.
  rcond.npz PartialPivLU::rcond / inverse (PartialPivLU.h:184-202) and
            LLT::rcond (LLT.h:161-166) on the lu.npz / llt.npz inputs
            (`make_golden.py rcond` writes this file alone).
  rankupdate.npz  LLT rank-one updates (internal::llt_rank_update_lower,
            LLT.h:231-299) of Eigen's factor of A = spd(seed, n): the updated
            lower factor and info for sigma = +2 and -0.5, and a downdate
            that fails at column 3 (`make_golden.py rankupdate`).
  ldlt.npz  LDLT<MatrixXd, Lower> (internal::ldlt_inplace<Lower>::unblocked,
            LDLT.h:290-395, and solve, :559-593) of symmetric indefinite,
            positive and negative definite matrices: transpositions, sign,
            info, the factor (n <= 24; diagonal, row and column sums else) and
            the solution of A x = b; plus small exact semidefinite and failing
            cases (`make_golden.py ldlt`).
  ldltrank.npz  LDLT<MatrixXd, Lower>::rankUpdate (LDLT.h:403-449, 529-555)
            of Eigen's LDLT of the ldlt.npz matrices (n = 5, 24, 64) by a
            seeded w with sigma = +1 and -0.5: the updated factor (n <= 24;
            diagonal, row and column sums else) and transpositions
            (`make_golden.py ldltrank`).
Inputs are not stored (they are regenerated from the seeds); their SHA-256
is, so a diverging generator is detected.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import pyoracle as po  # noqa: E402
import inputs  # noqa: E402

LU_SIZES = [1, 2, 3, 4, 5, 7, 8, 15, 16, 17, 31, 32, 33, 64, 127, 128, 129, 255, 256, 257]
LLT_SIZES = [1, 6, 9, 12, 15, 18, 24, 31, 32, 33, 127, 128, 129]
GEMM_SHAPES = [(8, 8, 8), (9, 9, 9), (24, 24, 24), (239, 239, 239), (240, 240, 240),
               (240, 24, 24), (24, 240, 24), (24, 24, 240), (24, 240, 240), (240, 24, 240),
               (240, 240, 24), (240, 240, 64), (240, 23, 160), (23, 240, 160)]


def lu_case(ref, a, b):
    n = a.shape[0]
    lu_ = np.asfortranarray(a.copy())
    t = np.zeros(n, dtype=np.int32)
    nbt = C.c_int(0)
    fz = ref.eref_lu_blocked_d(n, n, lu_, n, t, C.byref(nbt))
    x = np.zeros((n, 1), order="F")
    lu2 = np.zeros((n, n), order="F")
    perm = np.zeros(n, dtype=np.int32)
    det = C.c_double(0)
    ref.eref_partialpivlu_d(n, np.asfortranarray(a), n, 1, np.asfortranarray(b.reshape(n, 1)), n,
                            lu2, perm, x, n, C.byref(det))
    return lu_, t, nbt.value, fz, x[:, 0], perm, det.value


def main():
    ref = po.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libeigen_ref.so not built (needs the reference headers)")
    ref.eref_set_threads(1)
    out = {"eigen_version": np.array(ref.eref_version())}
    for n in LU_SIZES:
        a = inputs.normal(1000 + n, n, n)
        b = inputs.normal(2000 + n, n)
        lu_, t, nbt, fz, x, perm, det = lu_case(ref, a, b)
        out[f"n{n}_sha"] = np.array(inputs.sha256(a))
        out[f"n{n}_transp"] = t
        out[f"n{n}_perm"] = perm
        out[f"n{n}_nbt"] = np.array(nbt)
        out[f"n{n}_fz"] = np.array(fz)
        out[f"n{n}_x"] = x
        out[f"n{n}_det"] = np.array(det)
        if n <= 64:
            out[f"n{n}_lu"] = lu_
        else:
            out[f"n{n}_lu_fro"] = np.array(np.linalg.norm(lu_))
    # the plain FP64 drop-in configuration: n = 1024, seeds 1 / 2
    a = inputs.normal(1, 1024, 1024)
    b = inputs.normal(2, 1024)
    lu_, t, nbt, fz, x, perm, det = lu_case(ref, a, b)
    out["c1_sha_a"] = np.array(inputs.sha256(a))
    out["c1_sha_b"] = np.array(inputs.sha256(b))
    out["c1_transp"] = t
    out["c1_nbt"] = np.array(nbt)
    out["c1_x"] = x
    out["c1_berr"] = np.array(po.backward_error(a, x, b))
    # special cases
    z = inputs.normal(77, 40, 40)
    z[:, 5] = 0.0  # exact zero column -> first zero pivot, factorization continues
    lu_, t, nbt, fz, x, perm, det = lu_case(ref, z, inputs.normal(78, 40))
    out["zero_transp"] = t
    out["zero_fz"] = np.array(fz)
    out["zero_lu"] = lu_
    tie = np.asfortranarray(np.array([[1.0, 2.0, 0.0], [-1.0, 1.0, 3.0], [1.0, 0.0, 1.0]]))
    lu_, t, nbt, fz, x, perm, det = lu_case(ref, tie, np.ones(3))
    out["tie_transp"] = t
    out["tie_lu"] = lu_
    np.savez_compressed(os.path.join(HERE, "lu.npz"), **out)

    out = {}
    for n in LLT_SIZES:
        a = inputs.spd(3000 + n, n)
        b = inputs.normal(4000 + n, n)
        L = np.zeros((n, n), order="F")
        x = np.zeros((n, 1), order="F")
        info = ref.eref_llt_d(n, a, n, 1, np.asfortranarray(b.reshape(n, 1)), n, L, x, n)
        out[f"n{n}_sha"] = np.array(inputs.sha256(a))
        out[f"n{n}_info"] = np.array(info)
        out[f"n{n}_x"] = x[:, 0]
        if n <= 33:
            out[f"n{n}_L"] = np.tril(L)
    for n, bad in [(10, 4), (40, 35), (200, 150)]:
        a = inputs.spd(5000 + n, n)
        a[bad, bad] = -a[bad, bad] - 1e3  # not SPD from column `bad` on
        w = np.asfortranarray(a.copy())
        out[f"bad{n}_fail"] = np.array(ref.eref_llt_blocked_d(n, w, n))
        out[f"bad{n}_bad"] = np.array(bad)
    np.savez_compressed(os.path.join(HERE, "llt.npz"), **out)

    out = {}
    for (m, n, k) in GEMM_SHAPES:
        a = inputs.normal(6000 + m + 7 * n + 13 * k, m, k)
        b = inputs.normal(7000 + m + 7 * n + 13 * k, k, n)
        c = np.zeros((m, n), order="F")
        ref.eref_gemm_d(m, n, k, a, m, b, k, c, m)
        if m * n <= 24 * 240:
            out[f"g{m}_{n}_{k}"] = c
        else:  # row/column sums and a fixed sample pin orientation and values
            out[f"g{m}_{n}_{k}_rows"] = c.sum(axis=1)
            out[f"g{m}_{n}_{k}_cols"] = c.sum(axis=0)
            idx = inputs.rng(99).integers(0, m * n, size=256)
            out[f"g{m}_{n}_{k}_idx"] = idx
            out[f"g{m}_{n}_{k}_val"] = c.ravel(order="F")[idx]
    np.savez_compressed(os.path.join(HERE, "gemm.npz"), **out)

    out = {}
    n, offA, A, offb, b = inputs.packed_systems(11, 1000, [6, 9, 12, 15, 18, 24], np.float32, True)
    x = b.copy()
    info = np.zeros(1000, dtype=np.int32)
    ref.eref_batched_llt_solve_f(1000, offA, n, A, offb, x, info)
    out["llt_sha"] = np.array(inputs.sha256(A))
    out["llt_x"] = x
    out["llt_info"] = info
    n, offA, A, offB, B = inputs.packed_systems(12, 48, (16, 256), np.float64, False, nrhs=4)
    X = B.copy()
    info = np.zeros(48, dtype=np.int32)
    ref.eref_batched_lu_solve_d(48, offA, n, A, offB, 4, X, info)
    out["lu_sha"] = np.array(inputs.sha256(A))
    out["lu_x"] = X
    np.savez_compressed(os.path.join(HERE, "batched.npz"), **out)
    print("fixtures written to", HERE)


RCOND_LU_SIZES = [1, 2, 3, 5, 16, 33, 64, 129, 257]
RCOND_LLT_SIZES = [1, 6, 24, 33, 129]
INV_SIZES = [1, 5, 33, 64]


def rcond_main():
    ref = po.ref()
    out = {}
    for n in RCOND_LU_SIZES:
        a = inputs.normal(1000 + n, n, n)
        out[f"lu{n}_sha"] = np.array(inputs.sha256(a))
        out[f"lu{n}_rcond"] = np.array(ref.eref_lu_rcond_d(n, a, n))
    for n in RCOND_LLT_SIZES:
        a = inputs.spd(3000 + n, n)
        out[f"llt{n}_sha"] = np.array(inputs.sha256(a))
        out[f"llt{n}_rcond"] = np.array(ref.eref_llt_rcond_d(n, a, n))
    for n in INV_SIZES:
        a = inputs.normal(1000 + n, n, n)
        inv = np.zeros((n, n), order="F")
        ref.eref_lu_inverse_d(n, a, n, inv)
        out[f"inv{n}"] = inv
    np.savez_compressed(os.path.join(HERE, "rcond.npz"), **out)
    print("rcond fixtures written to", HERE)


RANKUPD_SIZES = [1, 6, 24, 33, 100, 256]
RANKUPD_SIGMAS = [("p", 2.0), ("m", -0.5)]


def rankupdate_identity_case():
    """(L, v) of the identity-rotation fixture: n = 6, v_0 = 0, v_3 = inf."""
    n = 6
    a = inputs.spd(7000 + n, n)
    l0 = np.tril(np.linalg.cholesky(a))
    v = inputs.normal(7100 + n, n)
    v[0] = 0.0
    v[3] = np.inf
    return np.asfortranarray(l0), v


def rankupdate_main():
    ref = po.ref()
    out = {}
    for n in RANKUPD_SIZES:
        a = inputs.spd(7000 + n, n)
        v = inputs.normal(7100 + n, n)
        out[f"n{n}_sha"] = np.array(inputs.sha256(a) + inputs.sha256(v))
        l0 = np.zeros((n, n), order="F")
        ref.eref_llt_d(n, a, n, 0, a, n, l0, a, n)
        for tag, sigma in RANKUPD_SIGMAS:
            f, info = po.llt_rank_update(l0, v, sigma, reference=True)
            lf = np.tril(f)
            if n <= 100:
                out[f"n{n}_{tag}_L"] = lf
            else:  # diagonal, row and column sums
                out[f"n{n}_{tag}_diag"] = np.diag(lf).copy()
                out[f"n{n}_{tag}_rows"] = lf.sum(axis=1)
                out[f"n{n}_{tag}_cols"] = lf.sum(axis=0)
            out[f"n{n}_{tag}_info"] = np.array(info)
    # a downdate that stops at column 3
    n = 24
    a = inputs.spd(7000 + n, n)
    l0 = np.zeros((n, n), order="F")
    ref.eref_llt_d(n, a, n, 0, a, n, l0, a, n)
    v = np.zeros(n)
    v[3] = 10.0 * np.sqrt(a[3, 3])
    _, info = po.llt_rank_update(l0, v, -1.0, reference=True)
    out["fail_info"] = np.array(info)
    # an identity rotation (v_0 = 0 over a positive diagonal) followed by a
    # non-finite entry in v: Eigen leaves column 0 untouched
    f, info = po.llt_rank_update(*rankupdate_identity_case(), 2.0, reference=True)
    out["ident_L"] = np.tril(f)
    out["ident_info"] = np.array(info)
    np.savez_compressed(os.path.join(HERE, "rankupdate.npz"), **out)
    print("rank update fixtures written to", HERE)


LDLT_SIZES = [1, 2, 5, 16, 24, 33, 64]
LDLT_KINDS = ["sym", "pos", "neg"]


LDLT_EXACT = {
    "zero3": np.zeros((3, 3)),
    "offdiag2": np.array([[0.0, 1.0], [1.0, 0.0]]),
    "midzero3": np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 1.0, 0.0]]),
    "psd4": np.diag([0.0, 3.0, 0.0, 1.0]),
    "nsd3": np.diag([-1.0, 0.0, -4.0]),
    "indef4": np.diag([2.0, -5.0, 0.0, 1.0]),
}


def ldlt_main():
    out = {}
    for kind in LDLT_KINDS:
        for n in LDLT_SIZES:
            a = inputs.ldlt_matrix(kind, n)
            b = inputs.normal(7600 + n, n)
            key = f"{kind}{n}"
            out[key + "_sha"] = np.array(inputs.sha256(a) + inputs.sha256(b))
            f, t, sg, ok = po.ldlt(a, reference=True)
            lf = np.tril(f)
            if n <= 24:
                out[key + "_f"] = lf
            else:
                out[key + "_diag"] = np.diag(lf).copy()
                out[key + "_rows"] = lf.sum(axis=1)
                out[key + "_cols"] = lf.sum(axis=0)
            out[key + "_t"] = t.copy()
            out[key + "_sign"] = np.array(sg)
            out[key + "_ok"] = np.array(ok)
            out[key + "_x"] = po.ldlt_solve_reference(a, b)
    for name, a in LDLT_EXACT.items():
        f, t, sg, ok = po.ldlt(np.asfortranarray(a), reference=True)
        out[name + "_f"] = np.tril(f)
        out[name + "_t"] = t.copy()
        out[name + "_sign"] = np.array(sg)
        out[name + "_ok"] = np.array(ok)
    np.savez_compressed(os.path.join(HERE, "ldlt.npz"), **out)
    print("ldlt fixtures written to", HERE)


LDLT_RANK_SIZES = [5, 24, 64]
LDLT_RANK_SIGMAS = {"p": 1.0, "m": -0.5}


def ldltrank_main():
    out = {}
    for kind in LDLT_KINDS:
        for n in LDLT_RANK_SIZES:
            a = inputs.ldlt_matrix(kind, n)
            w = inputs.normal(9000 + n, n)
            for tag, sigma in LDLT_RANK_SIGMAS.items():
                key = f"{kind}{n}{tag}"
                out[key + "_sha"] = np.array(inputs.sha256(a) + inputs.sha256(w))
                f, t = po.ldlt_rank_update_reference(a, w, sigma)
                lf = np.tril(f)
                if n <= 24:
                    out[key + "_f"] = lf
                else:
                    out[key + "_diag"] = np.diag(lf).copy()
                    out[key + "_rows"] = lf.sum(axis=1)
                    out[key + "_cols"] = lf.sum(axis=0)
                out[key + "_t"] = t.copy()
    np.savez_compressed(os.path.join(HERE, "ldltrank.npz"), **out)
    print("ldlt rank-update fixtures written to", HERE)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "rcond":
        rcond_main()
    elif len(sys.argv) > 1 and sys.argv[1] == "rankupdate":
        rankupdate_main()
    elif len(sys.argv) > 1 and sys.argv[1] == "ldlt":
        ldlt_main()
    elif len(sys.argv) > 1 and sys.argv[1] == "ldltrank":
        ldltrank_main()
    else:
        main()
