"""GPU parity tests of the FP64 drop-in path (GEMM, PartialPivLU, LLT,
triangular solves) against the oracle and the reference fixtures, through
the C ABI.  Pattern of the reference's differential CPU-vs-GPU tests
(test/cuda_common.h:65-80): same seeded input on both sides."""
import itertools
import os

import numpy as np
import pytest
import torch

import eigen_amd as ea
from oracle import pyoracle as po
from tests import inputs

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(__file__), "golden")
EPS = np.finfo(np.float64).eps


def dev(a):
    return ea.colmajor(torch.from_numpy(np.asfortranarray(a)))


def host(t):
    return t.cpu().numpy()


def first_pivot_difference(t_gpu, t_ref, margin):
    """Index of the first differing transposition, None if none; a difference
    at a near-tie (runner-up within 1e-12 relative) is excused."""
    d = np.nonzero(np.asarray(t_gpu) != np.asarray(t_ref))[0]
    if len(d) == 0:
        return None
    k = int(d[0])
    assert margin[k] < 1e-12, f"pivot {k}: gpu {t_gpu[k]} vs ref {t_ref[k]}, margin {margin[k]}"
    return k


# ----------------------------------------------------------------- GEMM --
@pytest.mark.parametrize("m,n,k", [(8, 8, 8), (9, 9, 9), (24, 24, 24), (239, 239, 239),
                                   (240, 240, 240), (2400, 24, 24), (24, 2400, 24),
                                   (24, 24, 2400), (24, 2400, 2400), (2400, 24, 2400),
                                   (2400, 2400, 24), (2400, 2400, 64), (4800, 23, 160),
                                   (23, 4800, 160), (1, 1, 1), (129, 130, 131)])
def test_dgemm_shapes(gpu, m, n, k):
    a = inputs.normal(6000 + m + 7 * n + 13 * k, m, k)
    b = inputs.normal(7000 + m + 7 * n + 13 * k, k, n)
    c = host(ea.gemm(dev(a), dev(b)))
    ref = a @ b
    bound = 2 * k * EPS * (np.abs(a) @ np.abs(b)) + 1e-300
    assert np.all(np.abs(c - ref) <= bound)


@pytest.mark.parametrize("alpha,beta", [(-1.5, 0.25), (-1.0, 1.0), (1.0, 1.0)])
@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_dgemm_transposes_alpha_beta(gpu, ta, tb, alpha, beta):
    """All transposes; beta = 1 with alpha = -1 or 1 takes the path whose
    accumulators start from C (the factorizations' trailing updates)."""
    m, n, k = 300, 200, 150
    a = inputs.normal(1, *(m, k) if ta == "N" else (k, m))
    b = inputs.normal(2, *(k, n) if tb == "N" else (n, k))
    c0 = inputs.normal(3, m, n)
    out = host(ea.gemm(dev(a), dev(b), dev(c0), alpha=alpha, beta=beta, transa=ta, transb=tb))
    opa = a if ta == "N" else a.T
    opb = b if tb == "N" else b.T
    np.testing.assert_allclose(out, alpha * opa @ opb + beta * c0, atol=1e-12)


def test_dgemm_reference_fixture(gpu):
    g = np.load(os.path.join(G, "gemm.npz"), allow_pickle=False)
    for (m, n, k) in [(8, 8, 8), (24, 240, 24), (240, 240, 64)]:
        a = inputs.normal(6000 + m + 7 * n + 13 * k, m, k)
        b = inputs.normal(7000 + m + 7 * n + 13 * k, k, n)
        c = host(ea.gemm(dev(a), dev(b)))
        key = f"g{m}_{n}_{k}"
        if key in g:
            np.testing.assert_allclose(c, g[key], atol=1e-13 * k)
        else:
            np.testing.assert_allclose(c.ravel(order="F")[g[key + "_idx"]], g[key + "_val"],
                                       atol=1e-13 * k)


def test_dsyrk_lower_and_upper(gpu):
    a = inputs.normal(5, 300, 70)
    c0 = inputs.normal(6, 300, 300)
    for uplo in "LU":
        c = dev(c0)
        ea.rank_update(c, dev(a), alpha=-1.0, uplo=uplo)
        out = host(c)
        full = c0 - a @ a.T
        tri = np.tril if uplo == "L" else np.triu
        other = np.triu if uplo == "L" else np.tril
        np.testing.assert_allclose(tri(out), tri(full), atol=1e-12)
        np.testing.assert_array_equal(other(out, 1 if uplo == "L" else -1),
                                      other(c0, 1 if uplo == "L" else -1))


# ------------------------------------------------------------------- LU --
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 15, 16, 17, 31, 32, 33, 64, 127, 128, 129,
                               255, 256, 257])
def test_lu_sweep_against_reference_fixture(gpu, n):
    g = np.load(os.path.join(G, "lu.npz"), allow_pickle=False)
    a = inputs.normal(1000 + n, n, n)
    assert inputs.sha256(a) == str(g[f"n{n}_sha"])
    lu = ea.PartialPivLU(dev(a))
    t = host(lu.transpositionsP())
    _, _, _, _, margin = po.lu(a, margins=True)
    k = first_pivot_difference(t, g[f"n{n}_transp"], margin)
    if k is None:
        assert lu.nb_transpositions() == int(g[f"n{n}_nbt"])
        np.testing.assert_array_equal(host(lu.permutationP()), g[f"n{n}_perm"])
        if n <= 64:
            np.testing.assert_allclose(host(lu.matrixLU()), g[f"n{n}_lu"], atol=1e-12 * n)
        assert lu.determinant() == pytest.approx(float(g[f"n{n}_det"]), rel=1e-9)
    b = inputs.normal(2000 + n, n)
    x = host(lu.solve(dev(b).reshape(-1)))
    assert po.backward_error(a, x, b) <= 1e-14
    rec = host(lu.reconstructedMatrix())
    np.testing.assert_allclose(rec, a, atol=1e-12 * max(1, n))


def test_lu_plain_fp64_config_n1024(gpu):
    """The plain drop-in configuration: n = 1024 N(0,1), seeded; pivots and
    backward error compared with Eigen's (fixture) and the oracle."""
    g = np.load(os.path.join(G, "lu.npz"), allow_pickle=False)
    a = inputs.normal(1, 1024, 1024)
    b = inputs.normal(2, 1024)
    lu = ea.PartialPivLU(dev(a))
    t = host(lu.transpositionsP())
    _, _, _, _, margin = po.lu(a, margins=True)
    k = first_pivot_difference(t, g["c1_transp"], margin)
    assert k is None, "no near-ties at this seed: pivots must be identical"
    x = host(lu.solve(dev(b).reshape(-1)))
    be = po.backward_error(a, x, b)
    assert be <= 1e-14
    # device-side backward error agrees with the long-double host one
    np.testing.assert_allclose(x, g["c1_x"], rtol=1e-8, atol=1e-10)


def test_lu_zero_pivot_and_tie_rules(gpu):
    g = np.load(os.path.join(G, "lu.npz"), allow_pickle=False)
    z = inputs.normal(77, 40, 40)
    z[:, 5] = 0.0
    lu = ea.PartialPivLU(dev(z))
    assert lu.first_zero_pivot() == 5
    np.testing.assert_array_equal(host(lu.transpositionsP()), g["zero_transp"])
    np.testing.assert_allclose(host(lu.matrixLU()), g["zero_lu"], atol=1e-12)
    tie = np.array([[1.0, 2.0, 0.0], [-1.0, 1.0, 3.0], [1.0, 0.0, 1.0]])
    lu = ea.PartialPivLU(dev(tie))
    np.testing.assert_array_equal(host(lu.transpositionsP()), g["tie_transp"])
    np.testing.assert_allclose(host(lu.matrixLU()), g["tie_lu"])


def test_lu_nan_rule(gpu):
    lu = ea.PartialPivLU(dev(np.array([[1.0, 0.0], [np.nan, 1.0]])))
    assert host(lu.transpositionsP())[0] == 0
    lu = ea.PartialPivLU(dev(np.array([[np.nan, 0.0], [5.0, 1.0]])))
    assert host(lu.transpositionsP())[0] == 0


@pytest.mark.parametrize("n", [600, 2048])
def test_lu_multi_workgroup_panel(gpu, n):
    """Sizes whose panels span several workgroups (grid-wide pivot exchange)."""
    a = inputs.normal(31 + n, n, n)
    lu = ea.PartialPivLU(dev(a))
    t = host(lu.transpositionsP())
    _, t_ref, _, _, margin = po.lu(a, margins=True)
    first_pivot_difference(t, t_ref, margin)
    b = inputs.normal(32 + n, n, 3)
    x = host(lu.solve(dev(b)))
    assert po.backward_error(a, x, b) <= 1e-14
    xt = host(lu.solve(dev(b), trans="T"))
    assert po.backward_error(np.asfortranarray(a.T), xt, b) <= 1e-14


def test_lu_host_buffer_entry_point(gpu):
    import ctypes as C
    n = 100
    a = inputs.normal(4, n, n)
    lu_h = np.asfortranarray(a.copy())
    t = np.zeros(n, dtype=np.int32)
    nbt, fz = C.c_int32(), C.c_int64()
    st = ea.lib().eamd_host_dgetrf(ea.handle(), n, n, lu_h.ctypes.data, n, t.ctypes.data,
                                   C.byref(nbt), C.byref(fz))
    assert st == 0
    lu_o, t_o, nbt_o, fz_o, margin = po.lu(a, margins=True)
    first_pivot_difference(t, t_o, margin)
    np.testing.assert_allclose(lu_h, lu_o, atol=1e-11)


# ------------------------------------------------------------------ LLT --
@pytest.mark.parametrize("n", [1, 6, 9, 24, 31, 32, 33, 127, 128, 129, 300, 1000])
def test_llt_against_oracle(gpu, n):
    a = inputs.spd(3000 + n, n)
    llt = ea.LLT(dev(a))
    assert llt.info() == ea.Success
    L = np.tril(host(llt.matrixLLT()))
    Lo, f = po.llt(a)
    np.testing.assert_allclose(L, np.tril(Lo), rtol=1e-10, atol=1e-10 * np.sqrt(n))
    # strictly upper part untouched, as the reference leaves it
    np.testing.assert_array_equal(np.triu(host(llt.matrixLLT()), 1), np.triu(a, 1))
    b = inputs.normal(4000 + n, n, 2)
    x = host(llt.solve(dev(b)))
    assert po.backward_error(a, x, b) <= 1e-14


@pytest.mark.parametrize("kappa", [1e6, 1e12])
def test_llt_ill_conditioned(gpu, kappa):
    """A = Q diag(lambda) Q^T with lambda log-spaced from 1 to 1/kappa (n =
    1000, eight block columns): the factor's residual ||A - L L^T||_F /
    ||A||_F stays within 4x the oracle's (a backward-stable Cholesky reaches
    a few 1e-16 whatever kappa), and the solve's backward error <= 1e-14."""
    n = 1000
    rng = np.random.default_rng(int(np.log10(kappa)))
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = (q * np.logspace(0, -np.log10(kappa), n)) @ q.T
    a = np.asfortranarray((a + a.T) / 2)
    llt = ea.LLT(dev(a))
    assert llt.info() == ea.Success
    L = np.tril(host(llt.matrixLLT()))
    Lo, _ = po.llt(a)
    Lo = np.tril(Lo)
    res = np.linalg.norm(a - L @ L.T) / np.linalg.norm(a)
    res_o = np.linalg.norm(a - Lo @ Lo.T) / np.linalg.norm(a)
    assert res <= max(4 * res_o, 1e-15), (res, res_o)
    b = inputs.normal(4100, n)
    assert po.backward_error(a, host(llt.solve(dev(b))), b) <= 1e-14


@pytest.mark.parametrize("nrhs", [5, 16, 21])
def test_solves_many_rhs(gpu, nrhs):
    """More than four right-hand sides take the sixteen-column triangular
    chain (21 = 16 + 5 also takes the four-column one): LU (both
    transpositions) and LLT solves meet the FP64 bound column by column."""
    n = 700
    a = inputs.normal(5000 + nrhs, n, n)
    b = inputs.normal(5100 + nrhs, n, nrhs)
    lu = ea.PartialPivLU(dev(a))
    x = host(lu.solve(dev(b)))
    assert po.backward_error(a, x, b) <= 1e-14
    xt = host(lu.solve(dev(b), trans="T"))
    assert po.backward_error(np.asfortranarray(a.T), xt, b) <= 1e-14
    s = inputs.spd(5200 + nrhs, n)
    llt = ea.LLT(dev(s))
    y = host(llt.solve(dev(b)))
    assert po.backward_error(s, y, b) <= 1e-14
    np.testing.assert_allclose(y, np.linalg.solve(s, b), rtol=1e-9, atol=1e-12)


def test_llt_upper(gpu):
    n = 200
    a = inputs.spd(9, n)
    llt = ea.LLT(dev(a), uplo="U")
    assert llt.info() == ea.Success
    U = np.triu(host(llt.matrixLLT()))
    np.testing.assert_allclose(U.T @ U, a, atol=1e-10 * n)
    b = inputs.normal(10, n)
    x = host(llt.solve(dev(b).reshape(-1)))
    assert po.backward_error(a, x, b) <= 1e-14


@pytest.mark.parametrize("n", [10, 40, 200])
def test_llt_failing_column(gpu, n):
    g = np.load(os.path.join(G, "llt.npz"), allow_pickle=False)
    a = inputs.spd(5000 + n, n)
    bad = int(g[f"bad{n}_bad"])
    a[bad, bad] = -a[bad, bad] - 1e3
    llt = ea.LLT(dev(a))
    assert llt.info() == ea.NumericalIssue
    assert llt.fail_column() == int(g[f"bad{n}_fail"])


# ----------------------------------------------------------------- TRSM --
@pytest.mark.parametrize("side,uplo,trans,diag",
                         list(itertools.product("LR", "LU", "NT", "NU")))
def test_trsm_all_variants(gpu, side, uplo, trans, diag):
    m, n = 150, 70
    na = m if side == "L" else n
    t = inputs.normal(11, na, na) + 4 * np.sqrt(na) * np.eye(na)
    t = np.asfortranarray(np.tril(t) if uplo == "L" else np.triu(t))
    b = inputs.normal(12, m, n)
    out = dev(b)
    ea.triangular_solve(dev(t), out, side=side, uplo=uplo, trans=trans, diag=diag, alpha=2.0)
    ref = po.trsm(side, uplo, trans, diag, t, b, alpha=2.0)
    np.testing.assert_allclose(host(out), ref, rtol=1e-10, atol=1e-11)


# ---------------------------------------------------------- rcond / inverse --
@pytest.fixture(scope="module")
def grc():
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "rcond.npz"))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 33, 64, 129, 257])
def test_lu_rcond_against_reference_fixture(gpu, grc, n):
    a = inputs.normal(1000 + n, n, n)
    lu = ea.dense.PartialPivLU(torch.from_numpy(a).cuda())
    np.testing.assert_allclose(lu.rcond(), float(grc[f"lu{n}_rcond"]), rtol=1e-9)


@pytest.mark.parametrize("n", [1, 6, 24, 33, 129])
def test_llt_rcond_against_reference_fixture(gpu, grc, n):
    a = inputs.spd(3000 + n, n)
    llt = ea.dense.LLT(torch.from_numpy(a).cuda())
    np.testing.assert_allclose(llt.rcond(), float(grc[f"llt{n}_rcond"]), rtol=1e-9)


@pytest.mark.parametrize("n", [1, 5, 33, 64])
def test_lu_inverse_against_reference_fixture(gpu, grc, n):
    a = inputs.normal(1000 + n, n, n)
    inv = ea.dense.PartialPivLU(torch.from_numpy(a).cuda()).inverse().cpu().numpy()
    ref = grc[f"inv{n}"]
    np.testing.assert_allclose(inv, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


def test_rcond_closed_forms_gpu(gpu):
    d = torch.tensor([3.0, -0.5, 2.0, 8.0, -1.0], dtype=torch.float64)
    assert ea.dense.PartialPivLU(torch.diag(d).cuda()).rcond() == pytest.approx(1 / 16, rel=1e-13)
    assert ea.dense.LLT(torch.diag(d.abs()).cuda()).rcond() == pytest.approx(1 / 16, rel=1e-13)
    z = torch.zeros(4, 4, dtype=torch.float64)
    assert ea.dense.PartialPivLU(z.cuda()).rcond() == 0.0


# ------------------------------------------------------- LLT rank update --
@pytest.mark.parametrize("n", [1, 6, 24, 33, 100, 256])
@pytest.mark.parametrize("tag,sigma", [("p", 2.0), ("m", -0.5)])
def test_llt_rank_update_against_reference_fixture(gpu, n, tag, sigma):
    g = np.load(os.path.join(G, "rankupdate.npz"), allow_pickle=False)
    a = inputs.spd(7000 + n, n)
    v = inputs.normal(7100 + n, n)
    llt = ea.LLT(dev(a))
    llt.rankUpdate(dev(v), sigma)
    assert llt.info() == ea.Success and int(g[f"n{n}_{tag}_info"]) == -1
    lf = np.tril(host(llt.matrixLLT()))
    if f"n{n}_{tag}_L" in g.files:
        np.testing.assert_allclose(lf, g[f"n{n}_{tag}_L"], rtol=1e-11, atol=1e-11)
    else:
        np.testing.assert_allclose(np.diag(lf), g[f"n{n}_{tag}_diag"], rtol=1e-11)
        np.testing.assert_allclose(lf.sum(axis=1), g[f"n{n}_{tag}_rows"], rtol=1e-9, atol=1e-9)
    target = a + sigma * np.outer(v, v)
    assert np.abs(lf @ lf.T - target).max() <= 1e-13 * np.abs(target).max()
    # the updated factor solves the updated system
    b = inputs.normal(7200 + n, n, 2)
    x = host(llt.solve(dev(b)))
    assert po.backward_error(target, x, b) <= 1e-14


def test_llt_rank_update_upper_and_failure(gpu):
    n = 24
    a = inputs.spd(7000 + n, n)
    v = inputs.normal(7100 + n, n)
    lo = ea.LLT(dev(a)).rankUpdate(dev(v), 2.0)
    up = ea.LLT(dev(a), uplo="U").rankUpdate(dev(v), 2.0)
    np.testing.assert_allclose(np.triu(host(up.matrixLLT())), np.tril(host(lo.matrixLLT())).T,
                               rtol=1e-13, atol=1e-14)
    g = np.load(os.path.join(G, "rankupdate.npz"), allow_pickle=False)
    w = np.zeros(n)
    w[3] = 10.0 * np.sqrt(a[3, 3])
    llt = ea.LLT(dev(a)).rankUpdate(dev(w), -1.0)
    assert llt.info() == ea.NumericalIssue and int(g["fail_info"]) == 3


def test_llt_rank_update_identity_rotation(gpu):
    """The identity rotation (v_0 = 0, L_00 > 0) is skipped as in Eigen, so
    an infinity in v reaches only the columns Eigen's update reaches."""
    from eigen_amd import batched
    from tests.golden.make_golden import rankupdate_identity_case
    g = np.load(os.path.join(G, "rankupdate.npz"), allow_pickle=False)
    l0, v = rankupdate_identity_case()
    n = l0.shape[0]
    L = dev(np.asfortranarray(l0).reshape(-1, order="F"))
    info = batched.llt_rank_update_batched(
        torch.zeros(1, dtype=torch.int64, device="cuda"),
        torch.full((1,), n, dtype=torch.int32, device="cuda"), L,
        torch.zeros(1, dtype=torch.int64, device="cuda"), dev(v),
        torch.full((1,), 2.0, dtype=torch.float64, device="cuda"))
    lf = host(L).reshape(n, n, order="F")
    assert int(info[0]) == int(g["ident_info"]) == -1
    # the non-finite entries where Eigen has them (assert_allclose checks
    # their positions and signs), the rest to rounding; column 0 untouched
    np.testing.assert_allclose(np.tril(lf), g["ident_L"], rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(lf[:, 0], l0[:, 0])


def test_llt_rank_update_batched(gpu):
    """Many small factors in one launch (the EKF-style workload): each
    against the oracle's update of the same GPU factor; a size above 256 is
    reported, not run."""
    from eigen_amd import batched
    rng = np.random.default_rng(11)
    sizes = rng.choice([6, 9, 12, 15, 18, 24], size=300)
    sizes[7] = 300
    offL = np.concatenate([[0], np.cumsum(sizes.astype(np.int64) ** 2)[:-1]])
    offv = np.concatenate([[0], np.cumsum(sizes.astype(np.int64))[:-1]])
    L = np.zeros(int((sizes.astype(np.int64) ** 2).sum()))
    V = rng.standard_normal(int(sizes.sum()))
    sig = np.where(rng.random(len(sizes)) < 0.5, 1.5, -0.25)
    for i, n in enumerate(sizes):
        if n > 256:
            continue
        a = inputs.spd(8000 + i, int(n))
        L[offL[i]:offL[i] + n * n] = np.linalg.cholesky(a).flatten(order="F")
    Ld = torch.tensor(L, device="cuda")
    info = batched.llt_rank_update_batched(
        torch.tensor(offL, device="cuda"), torch.tensor(sizes, dtype=torch.int32, device="cuda"),
        Ld, torch.tensor(offv, device="cuda"), torch.tensor(V, device="cuda"),
        torch.tensor(sig, device="cuda"))
    info = info.cpu().numpy()
    out = Ld.cpu().numpy()
    assert info[7] == -2
    for i, n in enumerate(sizes):
        if n > 256:
            continue
        l0 = L[offL[i]:offL[i] + n * n].reshape(n, n, order="F")
        ref, rinfo = po.llt_rank_update(l0, V[offv[i]:offv[i] + n], sig[i])
        assert info[i] == rinfo == -1
        got = out[offL[i]:offL[i] + n * n].reshape(n, n, order="F")
        np.testing.assert_allclose(np.tril(got), np.tril(ref), rtol=1e-12, atol=1e-12)


def test_fp32_single_matrix_entry_points(gpu):
    """eamd_sgetrf/sgetrs and eamd_spotrf/spotrs (FP32 storage, FP64
    arithmetic) through the C ABI on device buffers: the pivots are the
    oracle's FP64 pivots on the FP32 values, each factor entry is the FP64
    factor rounded once, both transposed and plain LU solves and the
    Cholesky solve reach FP32 backward error, and a non-SPD matrix reports
    its failing column."""
    import ctypes as C
    from eigen_amd import dense
    lib, h, p = dense.lib(), dense.handle(), dense._p
    n, nrhs = 300, 5
    A32 = inputs.normal(7100, n, n).astype(np.float32)
    A = A32.astype(np.float64)
    dA = ea.colmajor(torch.from_numpy(np.asfortranarray(A32)))
    t = torch.empty(n, dtype=torch.int32, device="cuda")
    nbt, fz = C.c_int32(0), C.c_int64(0)
    assert lib.eamd_sgetrf(h, n, n, p(dA), n, p(t), C.byref(nbt), C.byref(fz)) == 0
    lu_r, t_r, nbt_r, fz_r = po.lu(A)
    np.testing.assert_array_equal(host(t), t_r)
    assert (nbt.value, fz.value) == (nbt_r, fz_r)
    np.testing.assert_allclose(host(dA).astype(np.float64), lu_r, rtol=2.0 ** -24,
                               atol=1e-11 * np.abs(lu_r).max())
    b32 = inputs.normal(7101, n, nrhs).astype(np.float32)
    for trans in (b"N", b"T"):
        dB = ea.colmajor(torch.from_numpy(np.asfortranarray(b32)))
        assert lib.eamd_sgetrs(h, trans, n, nrhs, p(dA), n, p(t), p(dB), n) == 0
        op = A if trans == b"N" else A.T
        x = host(dB).astype(np.float64)
        assert po.backward_error(op, x, b32.astype(np.float64)) <= 1e-5
    S32 = inputs.spd(7102, n).astype(np.float32)
    S = S32.astype(np.float64)
    l_r, fail_r = po.llt(S)
    assert fail_r == -1
    for uplo in (b"L", b"U"):
        dS = ea.colmajor(torch.from_numpy(np.asfortranarray(S32)))
        fail = C.c_int64(0)
        assert lib.eamd_spotrf(h, uplo, n, p(dS), n, C.byref(fail)) == 0
        assert fail.value == -1
        F = host(dS).astype(np.float64)
        got, ref = (np.tril(F), np.tril(l_r)) if uplo == b"L" else (np.triu(F), np.tril(l_r).T)
        np.testing.assert_allclose(got, ref, rtol=2.0 ** -24, atol=1e-12 * np.abs(ref).max())
        # the other triangle keeps its input values exactly
        other = np.triu(F, 1) if uplo == b"L" else np.tril(F, -1)
        np.testing.assert_array_equal(other, np.triu(S, 1) if uplo == b"L" else np.tril(S, -1))
        dB = ea.colmajor(torch.from_numpy(np.asfortranarray(b32)))
        assert lib.eamd_spotrs(h, uplo, n, nrhs, p(dS), n, p(dB), n) == 0
        assert po.backward_error(S, host(dB).astype(np.float64), b32.astype(np.float64)) <= 1e-5
    bad = S32.copy()
    bad[40, 40] = -1.0
    dS = ea.colmajor(torch.from_numpy(np.asfortranarray(bad)))
    fail = C.c_int64(0)
    assert lib.eamd_spotrf(h, b"L", n, p(dS), n, C.byref(fail)) == 0
    assert fail.value == po.llt(bad.astype(np.float64))[1]
