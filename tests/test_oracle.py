"""CPU tests of the checker itself (oracle/oracle.c), pinned against
(1) the golden fixtures written from the reference (tests/golden/*.npz, made
by compiled Eigen), (2) closed forms and invariants, (3) brute force on tiny
inputs.  Each check is chosen so that a plausible mistake in the restatement
(dropped term, wrong sign or index, transposed operand) fails it."""
import itertools
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import inputs

G = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def glu():
    return np.load(os.path.join(G, "lu.npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def gllt():
    return np.load(os.path.join(G, "llt.npz"), allow_pickle=False)


# ------------------------------------------------------------------- LU --
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 15, 16, 17, 31, 32, 33, 64, 127, 128, 129,
                               255, 256, 257])
def test_lu_matches_reference_fixture(glu, n):
    a = inputs.normal(1000 + n, n, n)
    assert inputs.sha256(a) == str(glu[f"n{n}_sha"]), "input generator diverged"
    lu, t, nbt, fz, margin = po.lu(a, margins=True)
    ref_t = glu[f"n{n}_transp"]
    diff = np.nonzero(t != ref_t)[0]
    # identical pivots except at a near-tie (none expected at these seeds)
    if len(diff):
        assert margin[diff[0]] < 1e-12, f"pivot differs at column {diff[0]}"
    else:
        assert nbt == int(glu[f"n{n}_nbt"])
        assert fz == int(glu[f"n{n}_fz"])
    if n <= 64:
        np.testing.assert_allclose(lu, glu[f"n{n}_lu"], rtol=0, atol=1e-12 * max(1, n))
    else:
        assert abs(np.linalg.norm(lu) - float(glu[f"n{n}_lu_fro"])) <= 1e-10 * np.linalg.norm(lu)
    # permutation indices as PartialPivLU::permutationP() (PermutationMatrix.h:82-88)
    np.testing.assert_array_equal(po.perm_from_transpositions(t), glu[f"n{n}_perm"])
    b = inputs.normal(2000 + n, n)
    x = po.lu_solve(lu, t, b)[:, 0]
    np.testing.assert_allclose(x, glu[f"n{n}_x"], rtol=1e-9, atol=1e-9)
    # determinant = det_p * prod(diag U)
    det = (-1.0) ** nbt * np.prod(np.diag(lu))
    assert det == pytest.approx(float(glu[f"n{n}_det"]), rel=1e-10)


def test_lu_c1_instance_matches_reference(glu):
    a = inputs.normal(1, 1024, 1024)
    b = inputs.normal(2, 1024)
    assert inputs.sha256(a) == str(glu["c1_sha_a"])
    assert inputs.sha256(b) == str(glu["c1_sha_b"])
    lu, t, nbt, fz, margin = po.lu(a, margins=True)
    diff = np.nonzero(t != glu["c1_transp"])[0]
    assert len(diff) == 0 or margin[diff[0]] < 1e-12
    x = po.lu_solve(lu, t, b)[:, 0]
    be = po.backward_error(a, x, b)
    assert be < 1e-15
    assert float(glu["c1_berr"]) < 1e-15
    np.testing.assert_allclose(x, glu["c1_x"], rtol=1e-8, atol=1e-10)


def test_lu_reconstruction_pa_equals_lu():
    for n in (3, 20, 70, 200):
        a = inputs.normal(5 + n, n, n)
        lu, t, nbt, fz = po.lu(a)
        L = np.tril(lu, -1) + np.eye(n)
        U = np.triu(lu)
        perm = po.perm_from_transpositions(t)
        # (P A)[perm[i]] = A[i]
        pa = np.empty_like(a)
        pa[perm] = a
        np.testing.assert_allclose(L @ U, pa, atol=1e-12 * n)
        assert np.all(np.abs(L) <= 1.0 + 1e-15)  # partial pivoting bound


def test_lu_upper_triangular_input_has_identity_pivots():
    a = np.asfortranarray(np.triu(inputs.normal(9, 50, 50)) + 10 * np.eye(50))
    # make the diagonal dominant in its column so no interchange happens
    lu, t, nbt, fz = po.lu(a)
    np.testing.assert_array_equal(t, np.arange(50))
    assert nbt == 0
    np.testing.assert_allclose(lu, a)


def test_lu_permutation_matrix_input():
    n = 33
    p = inputs.rng(4).permutation(n)
    a = np.asfortranarray(np.eye(n)[p])
    lu, t, nbt, fz = po.lu(a)
    np.testing.assert_allclose(lu, np.eye(n))
    perm = po.perm_from_transpositions(t)
    pa = np.empty_like(a)
    pa[perm] = a
    np.testing.assert_allclose(pa, np.eye(n))


def test_lu_zero_column_reports_first_zero_pivot(glu):
    z = inputs.normal(77, 40, 40)
    z[:, 5] = 0.0
    lu, t, nbt, fz = po.lu(z)
    assert fz == int(glu["zero_fz"]) == 5
    np.testing.assert_array_equal(t, glu["zero_transp"])
    np.testing.assert_allclose(lu, glu["zero_lu"], atol=1e-12)


def test_lu_exact_tie_takes_lowest_row(glu):
    tie = np.asfortranarray(np.array([[1.0, 2.0, 0.0], [-1.0, 1.0, 3.0], [1.0, 0.0, 1.0]]))
    lu, t, nbt, fz = po.lu(tie)
    np.testing.assert_array_equal(t, glu["tie_transp"])
    assert t[0] == 0  # |1| == |-1| == |1|: first index wins (Visitor.h:179)
    np.testing.assert_allclose(lu, glu["tie_lu"])


def test_lu_nan_never_wins_below_the_diagonal():
    a = np.asfortranarray(np.array([[1.0, 0.0], [np.nan, 1.0]]))
    lu, t, nbt, fz = po.lu(a)
    assert t[0] == 0
    a = np.asfortranarray(np.array([[np.nan, 0.0], [5.0, 1.0]]))
    lu, t, nbt, fz = po.lu(a)
    assert t[0] == 0  # NaN in the first position is the scan's initial max


def test_transpositions_to_permutation_brute_force():
    for n in range(1, 6):
        for t in itertools.product(*[range(k, n) for k in range(n)]):
            t = np.array(t, dtype=np.int32)
            perm = po.perm_from_transpositions(t)
            # apply the transpositions in order to rows of the identity
            x = np.arange(n)
            for k in range(n):
                x[[k, t[k]]] = x[[t[k], k]]
            # P*b with b = arange: (P b)[perm[i]] = b[i]
            pb = np.empty(n, dtype=int)
            pb[perm] = np.arange(n)
            np.testing.assert_array_equal(pb, x)


# ------------------------------------------------------------------ LLT --
@pytest.mark.parametrize("n", [1, 6, 9, 12, 15, 18, 24, 31, 32, 33, 127, 128, 129])
def test_llt_matches_reference_fixture(gllt, n):
    a = inputs.spd(3000 + n, n)
    assert inputs.sha256(a) == str(gllt[f"n{n}_sha"])
    L, fail = po.llt(a)
    assert fail == -1 and int(gllt[f"n{n}_info"]) == 0
    if n <= 33:
        np.testing.assert_allclose(np.tril(L), gllt[f"n{n}_L"], rtol=1e-13, atol=1e-13)
    b = inputs.normal(4000 + n, n)
    x = po.llt_solve(L, b)[:, 0]
    np.testing.assert_allclose(x, gllt[f"n{n}_x"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(np.tril(L) @ np.tril(L).T, a, atol=1e-12 * n * np.abs(a).max())


@pytest.mark.parametrize("n", [10, 40, 200])
def test_llt_failing_column_matches_reference(gllt, n):
    a = inputs.spd(5000 + n, n)
    bad = int(gllt[f"bad{n}_bad"])
    a[bad, bad] = -a[bad, bad] - 1e3
    L, fail = po.llt(a)
    assert fail == int(gllt[f"bad{n}_fail"]) == bad


def test_llt_diagonal_closed_form():
    d = np.array([4.0, 9.0, 0.25, 16.0])
    L, fail = po.llt(np.diag(d))
    assert fail == -1
    np.testing.assert_allclose(np.diag(L), np.sqrt(d))
    L, fail = po.llt(np.diag([1.0, 2.0, -3.0, 4.0]))
    assert fail == 2


def test_llt_float_matches_double_within_single_precision():
    a = inputs.spd(8, 24)
    Ld, _ = po.llt(a)
    Lf, _ = po.llt(a.astype(np.float32), dtype=np.float32)
    np.testing.assert_allclose(np.tril(Lf), np.tril(Ld), rtol=2e-5, atol=2e-5)


# --------------------------------------------------------------- level 3 --
def test_gemm_matches_reference_fixture():
    g = np.load(os.path.join(G, "gemm.npz"), allow_pickle=False)
    shapes = [(8, 8, 8), (9, 9, 9), (24, 24, 24), (239, 239, 239), (240, 240, 240),
              (240, 24, 24), (24, 240, 24), (24, 24, 240), (24, 240, 240), (240, 24, 240),
              (240, 240, 24), (240, 240, 64), (240, 23, 160), (23, 240, 160)]
    for (m, n, k) in shapes:
        a = inputs.normal(6000 + m + 7 * n + 13 * k, m, k)
        b = inputs.normal(7000 + m + 7 * n + 13 * k, k, n)
        c = po.gemm(a, b)
        key = f"g{m}_{n}_{k}"
        tol = 1e-13 * k
        if key in g:
            np.testing.assert_allclose(c, g[key], atol=tol)
        else:
            np.testing.assert_allclose(c.sum(axis=1), g[key + "_rows"], atol=tol * n)
            np.testing.assert_allclose(c.sum(axis=0), g[key + "_cols"], atol=tol * m)
            np.testing.assert_allclose(c.ravel(order="F")[g[key + "_idx"]], g[key + "_val"],
                                       atol=tol)


def test_gemm_closed_forms_and_transposes():
    a = inputs.normal(1, 7, 5)
    b = inputs.normal(2, 5, 6)
    np.testing.assert_allclose(po.gemm(np.eye(7), a), a)
    np.testing.assert_allclose(po.gemm(a, b), a @ b, atol=1e-14)
    np.testing.assert_allclose(po.gemm(a.T, b, transa="T"), a @ b, atol=1e-14)
    np.testing.assert_allclose(po.gemm(a, b.T, transb="T"), a @ b, atol=1e-14)
    u = inputs.normal(3, 7)
    v = inputs.normal(4, 6)
    np.testing.assert_allclose(po.gemm(u.reshape(7, 1), v.reshape(1, 6)), np.outer(u, v))
    c0 = inputs.normal(5, 7, 6)
    np.testing.assert_allclose(po.gemm(a, b, c0, alpha=2.0, beta=-0.5), 2 * a @ b - 0.5 * c0,
                               atol=1e-13)


@pytest.mark.parametrize("side,uplo,trans,diag",
                         list(itertools.product("LR", "LU", "NT", "NU")))
def test_trsm_round_trip(side, uplo, trans, diag):
    m, n = 9, 7
    na = m if side == "L" else n
    t = inputs.normal(11, na, na) + 4 * np.eye(na)
    t = np.tril(t) if uplo == "L" else np.triu(t)
    top = t.copy()
    if diag == "U":
        np.fill_diagonal(top, 1.0)
    op = top.T if trans == "T" else top
    b = inputs.normal(12, m, n)
    x = po.trsm(side, uplo, trans, diag, t, b, alpha=1.5)
    lhs = op @ x if side == "L" else x @ op
    np.testing.assert_allclose(lhs, 1.5 * b, atol=1e-12)
    r = po.ref()
    if r is not None:  # the reference itself on the same call
        xr = np.asfortranarray(1.5 * b)
        r.eref_trsm_d(side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n,
                      np.asfortranarray(t), na, xr, m)
        np.testing.assert_allclose(x, xr, atol=1e-12)


def test_syrk_lower_equals_brute_force():
    a = inputs.normal(3, 10, 4)
    c = inputs.normal(4, 10, 10)
    out = po.syrk("L", "N", a, c, alpha=-1.0, beta=1.0)
    full = c - a @ a.T
    np.testing.assert_allclose(np.tril(out), np.tril(full), atol=1e-13)
    np.testing.assert_array_equal(np.triu(out, 1), np.triu(c, 1))


def test_backward_error_of_exact_solution_is_zero():
    a = np.asfortranarray(np.diag([2.0, 4.0, 8.0]))
    x = np.array([1.0, 1.0, 1.0])
    b = a @ x
    assert po.backward_error(a, x, b) == 0.0
    assert po.backward_error(a, x + np.array([1.0, 0, 0]), b) == pytest.approx(2.0 / (8 * 2 + 8))


# -------------------------------------------------------------- batched --
def test_batched_llt_sample_matches_reference():
    g = np.load(os.path.join(G, "batched.npz"), allow_pickle=False)
    n, offA, A, offb, b = inputs.packed_systems(11, 1000, [6, 9, 12, 15, 18, 24], np.float32,
                                                True)
    assert inputs.sha256(A) == str(g["llt_sha"])
    x, info = po.batched_llt_solve_f(offA, n, A, offb, b)
    assert np.all(info == -1) and np.all(g["llt_info"] == -1)
    np.testing.assert_allclose(x, g["llt_x"], rtol=1e-4, atol=1e-5)


def test_batched_lu_sample_matches_reference():
    g = np.load(os.path.join(G, "batched.npz"), allow_pickle=False)
    n, offA, A, offB, B = inputs.packed_systems(12, 48, (16, 256), np.float64, False, nrhs=4)
    assert inputs.sha256(A) == str(g["lu_sha"])
    X, info = po.batched_lu_solve_d(offA, n, A, offB, 4, B)
    np.testing.assert_allclose(X, g["lu_x"], rtol=1e-8, atol=1e-9)


def test_batched_equals_single_system_routines():
    n, offA, A, offb, b = inputs.packed_systems(13, 20, [6, 9, 24], np.float32, True)
    x, info = po.batched_llt_solve_f(offA, n, A, offb, b)
    for i in range(20):
        k = int(n[i])
        a = A[offA[i]:offA[i] + k * k].reshape(k, k, order="F")
        L, f = po.llt(a, dtype=np.float32)
        xi = po.llt_solve(L, b[offb[i]:offb[i] + k])[:, 0]
        np.testing.assert_array_equal(xi, x[offb[i]:offb[i] + k])


# ---------------------------------------------------------- rcond / inverse --
@pytest.fixture(scope="module")
def grc():
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "rcond.npz"))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 33, 64, 129, 257])
def test_lu_rcond_matches_reference_fixture(grc, n):
    a = inputs.normal(1000 + n, n, n)
    assert inputs.sha256(a) == str(grc[f"lu{n}_sha"])
    np.testing.assert_allclose(po.lu_rcond(a), float(grc[f"lu{n}_rcond"]), rtol=1e-10)


@pytest.mark.parametrize("n", [1, 6, 24, 33, 129])
def test_llt_rcond_matches_reference_fixture(grc, n):
    a = inputs.spd(3000 + n, n)
    assert inputs.sha256(a) == str(grc[f"llt{n}_sha"])
    np.testing.assert_allclose(po.llt_rcond(a), float(grc[f"llt{n}_rcond"]), rtol=1e-10)


def test_rcond_closed_forms():
    # identity: 1; diagonal: min|d| / max|d| (the estimator finds e_j exactly);
    # a zero matrix: 0
    assert po.lu_rcond(np.eye(7)) == pytest.approx(1.0)
    d = np.array([3.0, -0.5, 2.0, 8.0, -1.0])
    assert po.lu_rcond(np.diag(d)) == pytest.approx(0.5 / 8.0, rel=1e-14)
    assert po.llt_rcond(np.diag(np.abs(d))) == pytest.approx(0.5 / 8.0, rel=1e-14)
    assert po.rcond_from(0.0, 3, None, None) == 0.0
    # a 2x2 case by hand: A = [[1, 2], [3, 4]], ||A||_1 = 6, A^-1 = [[-2, 1], [1.5, -0.5]],
    # ||A^-1||_1 = 3.5
    assert po.lu_rcond(np.array([[1.0, 2.0], [3.0, 4.0]])) == pytest.approx(1.0 / 21.0, rel=1e-14)


def test_reference_inverse_fixture_is_an_inverse(grc):
    for n in (1, 5, 33, 64):
        a = inputs.normal(1000 + n, n, n)
        inv = grc[f"inv{n}"]
        assert np.abs(a @ inv - np.eye(n)).max() < 1e-10 * max(1.0, np.abs(inv).max() * n)


# ------------------------------------------------------- LLT rank update --
RANKUPD_SIZES = [1, 6, 24, 33, 100, 256]


@pytest.fixture(scope="module")
def grk():
    return np.load(os.path.join(G, "rankupdate.npz"), allow_pickle=False)


def _rk_check(g, key, lf):
    if f"{key}_L" in g.files:
        np.testing.assert_allclose(np.tril(lf), g[f"{key}_L"], rtol=1e-12, atol=1e-12)
    else:
        lf = np.tril(lf)
        np.testing.assert_allclose(np.diag(lf), g[f"{key}_diag"], rtol=1e-12)
        np.testing.assert_allclose(lf.sum(axis=1), g[f"{key}_rows"], rtol=1e-10, atol=1e-10)
        np.testing.assert_allclose(lf.sum(axis=0), g[f"{key}_cols"], rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("n", RANKUPD_SIZES)
@pytest.mark.parametrize("tag,sigma", [("p", 2.0), ("m", -0.5)])
def test_llt_rank_update_matches_reference_fixture(grk, n, tag, sigma):
    """The oracle's restated rank update (Givens for sigma > 0, the downdate
    recurrence otherwise) against Eigen's on the same factor, and the
    defining identity L' L'^T = A + sigma v v^T."""
    a = inputs.spd(7000 + n, n)
    v = inputs.normal(7100 + n, n)
    assert str(grk[f"n{n}_sha"]) == inputs.sha256(a) + inputs.sha256(v)
    l0, _ = po.llt(a)
    lf, info = po.llt_rank_update(np.tril(l0), v, sigma)
    assert info == int(grk[f"n{n}_{tag}_info"]) == -1
    _rk_check(grk, f"n{n}_{tag}", lf)
    lf = np.tril(lf)
    target = a + sigma * np.outer(v, v)
    assert np.abs(lf @ lf.T - target).max() <= 1e-13 * np.abs(target).max()


def test_llt_rank_downdate_failing_column(grk):
    n = 24
    a = inputs.spd(7000 + n, n)
    l0, _ = po.llt(a)
    v = np.zeros(n)
    v[3] = 10.0 * np.sqrt(a[3, 3])
    _, info = po.llt_rank_update(np.tril(l0), v, -1.0)
    assert info == int(grk["fail_info"]) == 3


def test_llt_rank_update_identity_rotation(grk):
    """v_0 = 0 over a positive L_00 makes the identity rotation, which
    Eigen does not apply (Jacobi.h:460-461): column 0 keeps its values
    although v holds an infinity further down.  The oracle matches Eigen's
    output entry for entry, infinities included."""
    from tests.golden.make_golden import rankupdate_identity_case
    l0, v = rankupdate_identity_case()
    lf, info = po.llt_rank_update(l0, v, 2.0)
    assert info == int(grk["ident_info"]) == -1
    # the non-finite entries where Eigen has them, the rest to rounding
    np.testing.assert_allclose(np.tril(lf), grk["ident_L"], rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(lf[:, 0], l0[:, 0])


def test_llt_rank_update_closed_form():
    """Diagonal A and v = e_k: only L_kk changes, to sqrt(A_kk + sigma)."""
    n = 7
    d = np.arange(1.0, n + 1.0)
    l0 = np.diag(np.sqrt(d))
    for sigma in (3.0, -0.5):
        v = np.zeros(n)
        v[4] = 1.0
        lf, info = po.llt_rank_update(l0, v, sigma)
        expect = np.sqrt(d.copy())
        expect[4] = np.sqrt(d[4] + sigma)
        assert info == -1
        np.testing.assert_allclose(np.diag(lf), expect, rtol=1e-15)
        np.testing.assert_array_equal(np.tril(lf, -1), 0.0)


# ------------------------------------------------------------------ LDLT --
LDLT_SIZES = [1, 2, 5, 16, 24, 33, 64]
LDLT_KINDS = ["sym", "pos", "neg"]
LDLT_EXACT = ["zero3", "offdiag2", "midzero3", "psd4", "nsd3", "indef4"]


@pytest.fixture(scope="module")
def gld():
    return np.load(os.path.join(G, "ldlt.npz"), allow_pickle=False)


def ldlt_check_factor(g, key, f, rtol):
    """The lower factor against the fixture: whole (n <= 24) or its
    diagonal, row and column sums."""
    lf = np.tril(f)
    scale = max(1.0, np.abs(lf).max())
    if f"{key}_f" in g.files:
        assert np.abs(lf - g[f"{key}_f"]).max() <= rtol * scale
    else:
        assert np.abs(np.diag(lf) - g[f"{key}_diag"]).max() <= rtol * scale
        n = lf.shape[0]
        assert np.abs(lf.sum(axis=1) - g[f"{key}_rows"]).max() <= rtol * scale * n
        assert np.abs(lf.sum(axis=0) - g[f"{key}_cols"]).max() <= rtol * scale * n


def ldlt_reconstruct(f, t):
    """P^T L D L^T P from a factor and its transpositions; (P e_i = e_s(i),
    s(i) following i through the transpositions in order)."""
    n = f.shape[0]
    L = np.tril(f, -1) + np.eye(n)
    res = (L * np.diag(f)) @ L.T
    s = []
    for i in range(n):
        pos = i
        for k, j in enumerate(t):
            if pos == k:
                pos = int(j)
            elif pos == j:
                pos = k
        s.append(pos)
    return res[np.ix_(s, s)]


@pytest.mark.parametrize("kind", LDLT_KINDS)
@pytest.mark.parametrize("n", LDLT_SIZES)
def test_ldlt_matches_reference_fixture(gld, kind, n):
    """The restated LDLT (pivot rule, lower-only exchange, left-looking
    column, sign and info) against Eigen's outputs on the same matrix, and
    the defining identity P^T L D L^T P = A."""
    a = inputs.ldlt_matrix(kind, n)
    b = inputs.normal(7600 + n, n)
    key = f"{kind}{n}"
    assert str(gld[key + "_sha"]) == inputs.sha256(a) + inputs.sha256(b)
    f, t, sg, ok = po.ldlt(a)
    np.testing.assert_array_equal(t, gld[key + "_t"])
    assert sg == int(gld[key + "_sign"]) and ok == bool(gld[key + "_ok"])
    ldlt_check_factor(gld, key, f, 1e-10 if kind == "sym" else 1e-13)
    assert np.abs(ldlt_reconstruct(f, t) - a).max() <= 1e-12 * max(1.0, np.abs(a).max()) * n
    x = po.ldlt_solve(f, t, b)
    xr = gld[key + "_x"]
    assert np.abs(x - xr).max() <= 1e-8 * max(1.0, np.abs(xr).max())
    assert np.abs(a @ x - b).max() <= 1e-9 * (np.abs(a).max() * np.abs(x).max() + 1.0) * n


@pytest.mark.parametrize("name", LDLT_EXACT)
def test_ldlt_exact_cases(gld, name):
    """Semidefinite, all-zero and failing inputs whose factors are exact:
    transpositions, sign, info and factor equal Eigen's bit for bit."""
    from tests.golden.make_golden import LDLT_EXACT as MATS
    f, t, sg, ok = po.ldlt(np.asfortranarray(MATS[name]))
    np.testing.assert_array_equal(t, gld[name + "_t"])
    assert sg == int(gld[name + "_sign"]) and ok == bool(gld[name + "_ok"])
    np.testing.assert_array_equal(np.tril(f), gld[name + "_f"])


def test_ldlt_closed_forms():
    """2 x 2 by hand, the pivot order of a diagonal matrix (largest |d|
    first, the lower index on ties) and the sign states."""
    a, b, c = 4.0, 2.0, -3.0
    f, t, sg, ok = po.ldlt(np.array([[a, b], [b, c]]))
    assert list(t) == [0, 1] and ok and sg == 3
    np.testing.assert_allclose([f[0, 0], f[1, 0], f[1, 1]], [a, b / a, c - b * b / a],
                               rtol=1e-15)
    f, t, sg, ok = po.ldlt(np.array([[c, b], [b, a]]))  # pivot 1 first
    assert list(t) == [1, 1] and ok
    np.testing.assert_allclose([f[0, 0], f[1, 0], f[1, 1]], [a, b / a, c - b * b / a],
                               rtol=1e-15)
    d = np.array([1.0, -7.0, 3.0, 7.0, -2.0])
    f, t, sg, ok = po.ldlt(np.diag(d))
    assert list(t) == [1, 3, 2, 4, 4] and ok and sg == 3
    np.testing.assert_array_equal(np.diag(f), [-7.0, 7.0, 3.0, -2.0, 1.0])
    assert po.ldlt(np.diag([2.0, 0.0, 1.0]))[2] == 0  # positive semidefinite
    assert po.ldlt(np.diag([-2.0, 0.0]))[2] == 1
    assert po.ldlt(np.zeros((4, 4)))[2:] == (2, True)
    assert po.ldlt(np.zeros((0, 0)))[2:] == (2, True)


def test_ldlt_brute_force_small():
    """Tiny random matrices: the factor is the unique unit-lower L and
    diagonal D with P A P^T = L D L^T for the recorded P, compared against
    a direct elimination of the permuted matrix."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        n = int(rng.integers(1, 6))
        g = rng.standard_normal((n, n))
        a = g + g.T
        f, t, sg, ok = po.ldlt(a)
        # apply the transpositions to A symmetrically, then eliminate without pivoting
        pa = a.copy()
        for k, j in enumerate(t):
            pa[[k, j], :] = pa[[j, k], :]
            pa[:, [k, j]] = pa[:, [j, k]]
        L = np.eye(n)
        D = np.zeros(n)
        w = pa.copy()
        for k in range(n):
            D[k] = w[k, k]
            L[k + 1:, k] = w[k + 1:, k] / D[k]
            w[k + 1:, k + 1:] -= np.outer(L[k + 1:, k], w[k, k + 1:])
        np.testing.assert_allclose(np.diag(f), D, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(np.tril(f, -1), np.tril(L, -1), rtol=1e-10, atol=1e-12)


LDLT_RANK_SIZES = [5, 24, 64]
LDLT_RANK_SIGMAS = {"p": 1.0, "m": -0.5}


@pytest.fixture(scope="module")
def gldr():
    return np.load(os.path.join(G, "ldltrank.npz"), allow_pickle=False)


@pytest.mark.parametrize("kind", LDLT_KINDS)
@pytest.mark.parametrize("n", LDLT_RANK_SIZES)
@pytest.mark.parametrize("tag", sorted(LDLT_RANK_SIGMAS))
def test_ldlt_rank_update_matches_reference_fixture(gldr, kind, n, tag):
    """The oracle's LDLT rank-one update (LDLT.h:403-449) of its own LDLT
    factor against Eigen's LDLT(A).rankUpdate(w, sigma) (transpositions
    equal, factor within 1e-10 / 1e-13 of its largest entry), and,
    independently of the reference, P^T L' D' L'^T P = A + sigma w w^T."""
    sigma = LDLT_RANK_SIGMAS[tag]
    a = inputs.ldlt_matrix(kind, n)
    w = inputs.normal(9000 + n, n)
    key = f"{kind}{n}{tag}"
    assert str(gldr[key + "_sha"]) == inputs.sha256(a) + inputs.sha256(w)
    f, t, _, _ = po.ldlt(a)
    fo = po.ldlt_rank_update(f, t, w, sigma)
    np.testing.assert_array_equal(t, gldr[key + "_t"])
    ldlt_check_factor(gldr, key, fo, 1e-10 if kind == "sym" else 1e-13)
    target = a + sigma * np.outer(w, w)
    err = np.abs(ldlt_reconstruct(fo, t) - target).max()
    assert err <= 1e-12 * max(1.0, np.abs(target).max()) * n
    np.testing.assert_array_equal(np.triu(fo, 1), np.triu(a, 1))
