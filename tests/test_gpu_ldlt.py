"""GPU parity of the batched LDLT (LDLT.h:290-395 compute, :559-593 solve)
against Eigen's outputs (tests/golden/ldlt.npz) and the oracle."""
import numpy as np
import pytest
import torch

import eigen_amd as ea
from eigen_amd import batched
from oracle import pyoracle as po
from tests import inputs
from tests.golden.make_golden import LDLT_EXACT
from tests.test_oracle import (LDLT_KINDS, LDLT_SIZES, ldlt_check_factor, ldlt_reconstruct)

pytestmark = pytest.mark.gpu

D64 = dict(dtype=torch.float64, device="cuda")


@pytest.fixture(scope="module")
def gld():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "ldlt.npz"),
                   allow_pickle=False)


def pack(mats):
    """Column-major packing of square matrices: (offsets, sizes, flat)."""
    sizes = [m.shape[0] for m in mats]
    off = np.concatenate([[0], np.cumsum([s * s for s in sizes])[:-1]]).astype(np.int64)
    flat = np.concatenate([np.asarray(m, dtype=np.float64).reshape(-1, order="F")
                           for m in mats] + [np.zeros(1)])
    return off, np.array(sizes, dtype=np.int32), flat


def run_ldlt(mats, max_n=None):
    off, sizes, flat = pack(mats)
    offA = torch.from_numpy(off).cuda()
    n = torch.from_numpy(sizes).cuda()
    A = torch.from_numpy(flat).cuda()
    T, offT, sign, info = batched.ldlt_batched(offA, n, A, max_n=max_n)
    torch.cuda.synchronize()
    return offA, n, A, T, offT, sign, info, off, sizes


def unpack(A, off, sizes, i):
    k = int(sizes[i])
    return A[int(off[i]):int(off[i]) + k * k].cpu().numpy().reshape(k, k, order="F")


def test_ldlt_batched_matches_reference_fixture(gld):
    """Every fixture matrix in one uneven batch: transpositions, sign and
    info equal Eigen's; the factor within 1e-10 (indefinite) / 1e-13
    (definite) of Eigen's, relative to its largest entry; the strict upper
    part untouched; the solve against Eigen's solution."""
    keys, mats, rhs = [], [], []
    for kind in LDLT_KINDS:
        for n in LDLT_SIZES:
            keys.append(f"{kind}{n}")
            mats.append(inputs.ldlt_matrix(kind, n))
            rhs.append(inputs.normal(7600 + n, n))
    offA, n, A, T, offT, sign, info, off, sizes = run_ldlt(mats)
    Th, sh, ih = T.cpu().numpy(), sign.cpu().numpy(), info.cpu().numpy()
    ot = offT.cpu().numpy()
    for i, key in enumerate(keys):
        k = int(sizes[i])
        print("ldlt fixture", key, flush=True)
        np.testing.assert_array_equal(Th[ot[i]:ot[i] + k], gld[key + "_t"])
        assert int(sh[i]) == int(gld[key + "_sign"])
        assert (int(ih[i]) == 0) == bool(gld[key + "_ok"]) and int(ih[i]) in (0, 1)
        f = unpack(A, off, sizes, i)
        ldlt_check_factor(gld, key, f, 1e-10 if key.startswith("sym") else 1e-13)
        np.testing.assert_array_equal(np.triu(f, 1), np.triu(mats[i], 1))
    offb = torch.from_numpy(np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)).cuda()
    b = torch.from_numpy(np.concatenate(rhs)).cuda()
    batched.ldlt_solve_batched(offA, n, A, offT, T, offb, b)
    bh = b.cpu().numpy()
    ob = offb.cpu().numpy()
    for i, key in enumerate(keys):
        k = int(sizes[i])
        x, xr = bh[ob[i]:ob[i] + k], gld[key + "_x"]
        assert np.abs(x - xr).max() <= 1e-8 * max(1.0, np.abs(xr).max()), key


def test_ldlt_exact_cases_bit_for_bit(gld):
    """Zero, semidefinite and failing matrices: factor, transpositions,
    sign and info exactly Eigen's."""
    names = list(LDLT_EXACT)
    mats = [LDLT_EXACT[k] for k in names]
    offA, n, A, T, offT, sign, info, off, sizes = run_ldlt(mats)
    Th, ot = T.cpu().numpy(), offT.cpu().numpy()
    for i, name in enumerate(names):
        k = int(sizes[i])
        np.testing.assert_array_equal(Th[ot[i]:ot[i] + k], gld[name + "_t"])
        assert int(sign[i]) == int(gld[name + "_sign"]), name
        assert (int(info[i]) == 0) == bool(gld[name + "_ok"]), name
        np.testing.assert_array_equal(np.tril(unpack(A, off, sizes, i)), gld[name + "_f"])


@pytest.mark.parametrize("max_n,count", [(64, 3000), (32, 2001), (16, 1001)])
def test_ldlt_many_uneven_systems_against_oracle(max_n, count):
    """Systems of sizes 0..max_n (indefinite, definite, semidefinite
    rank-deficient by an exact zero block; 1, 2 or 4 systems per wavefront,
    the last wavefront part-filled): every output against the oracle on the
    same input; P^T L D L^T P = A; then the solve against the oracle's."""
    rng = np.random.default_rng(11 + max_n)
    mats = []
    for i in range(count):
        k = int(rng.integers(0, max_n + 1))
        g = rng.standard_normal((k, k))
        kind = i % 4
        if kind == 0:
            a = g + g.T
        elif kind == 1:
            a = g @ g.T + k * np.eye(k)
        elif kind == 2:
            a = -(g @ g.T) - np.eye(k)
        else:
            a = np.zeros((k, k))
            h = k // 2
            a[:h, :h] = (g @ g.T)[:h, :h] + np.eye(h)
        mats.append(np.asfortranarray(a))
    print("ldlt uneven batch max_n", max_n, "count", count, flush=True)
    offA, n, A, T, offT, sign, info, off, sizes = run_ldlt(mats, max_n=max_n)
    Th, ot = T.cpu().numpy(), offT.cpu().numpy()
    sh, ih = sign.cpu().numpy(), info.cpu().numpy()
    Ah = A.cpu().numpy()
    for i, a in enumerate(mats):
        k = a.shape[0]
        f, t, sg, ok = po.ldlt(a)
        np.testing.assert_array_equal(Th[ot[i]:ot[i] + k], t, err_msg=str(i))
        assert int(sh[i]) == sg and (int(ih[i]) == 0) == ok, i
        if k == 0:
            continue
        fg = Ah[int(off[i]):int(off[i]) + k * k].reshape(k, k, order="F")
        scale = max(1.0, np.abs(np.tril(f)).max())
        # indefinite inputs: pivoting on the original diagonal lets entries
        # grow, and rounding differences grow with them (Eigen's own factor
        # differs from the oracle's by up to 8e-8 of its largest entry on
        # these inputs); definite ones stay near machine precision
        tol = 1e-6 if i % 4 == 0 else 1e-12
        assert np.abs(np.tril(fg) - np.tril(f)).max() <= tol * scale, i
        if i % 4 != 3:
            # backward error: Eigen's and the oracle's stay below 4e-13 here
            assert np.abs(ldlt_reconstruct(fg, t) - a).max() <= 1e-11 * max(1.0, np.abs(a).max()) * k
    # the solve with the GPU factors against the oracle's solve with them
    offb_h = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    bh = rng.standard_normal(int(sizes.sum()) + 1)
    b = torch.from_numpy(bh.copy()).cuda()
    batched.ldlt_solve_batched(offA, n, A, offT, T, torch.from_numpy(offb_h).cuda(), b,
                               max_n=max_n)
    xg = b.cpu().numpy()
    for i, a in enumerate(mats):
        k = a.shape[0]
        if k == 0 or i % 4 == 3:
            continue
        fg = Ah[int(off[i]):int(off[i]) + k * k].reshape(k, k, order="F")
        o = int(offb_h[i])
        x = po.ldlt_solve(fg, Th[ot[i]:ot[i] + k], bh[o:o + k])
        assert np.abs(xg[o:o + k] - x).max() <= 1e-10 * max(1.0, np.abs(x).max()), i


def test_ldlt_size_limits():
    """A size above max_n reports -2 and leaves its matrix alone; max_n
    above 64 is refused at the boundary."""
    a = np.eye(5)
    b = np.diag([1.0, 2.0, 3.0])
    offA, n, A, T, offT, sign, info, off, sizes = run_ldlt([a, b], max_n=4)
    assert int(info[0]) == -2 and int(info[1]) == 0
    np.testing.assert_array_equal(unpack(A, off, sizes, 0), a)
    # the solve with the same max_n leaves the -2 system's right-hand side
    # alone and solves the other one
    rhs = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 2.0, 4.0, 9.0])
    offb = torch.tensor([0, 5], dtype=torch.int64, device="cuda")
    bt = torch.from_numpy(rhs.copy()).cuda()
    batched.ldlt_solve_batched(offA, n, A, offT, T, offb, bt, max_n=4)
    bh = bt.cpu().numpy()
    np.testing.assert_array_equal(bh[:5], rhs[:5])
    np.testing.assert_allclose(bh[5:], [2.0, 2.0, 3.0], rtol=1e-15)
    big = np.eye(65)
    with pytest.raises(Exception):
        run_ldlt([big], max_n=65)


def test_ldlt_offsets_out_of_order():
    """Transposition offsets given in descending order: the buffer is
    sized by the furthest slot, and each system's transpositions and solve
    equal those of the same matrix factored alone."""
    mats = [inputs.ldlt_matrix("sym", 16), inputs.ldlt_matrix("pos", 5)]
    off, sizes, flat = pack(mats)
    offA = torch.from_numpy(off).cuda()
    n = torch.from_numpy(sizes).cuda()
    A = torch.from_numpy(flat).cuda()
    offT = torch.tensor([5, 0], dtype=torch.int64, device="cuda")
    T, _, sign, info = batched.ldlt_batched(offA, n, A, offT=offT)
    assert T.numel() >= 21
    Th = T.cpu().numpy()
    for i, m in enumerate(mats):
        _, t_ref, s_ref, ok = po.ldlt(m)
        o = int(offT[i])
        np.testing.assert_array_equal(Th[o:o + m.shape[0]], t_ref)
        assert int(sign[i]) == s_ref and (int(info[i]) == 0) == ok


def test_ldlt_class_mirror(gld):
    """LDLT(A): matrixLDLT / transpositionsP / vectorD / isPositive /
    isNegative / solve / reconstructedMatrix on the fixture's n = 24."""
    for kind in LDLT_KINDS:
        a = inputs.ldlt_matrix(kind, 24)
        b = inputs.normal(7624, 24)
        f = ea.LDLT(torch.from_numpy(a).cuda())
        key = f"{kind}24"
        assert f.info() == ea.Success
        np.testing.assert_array_equal(f.transpositionsP().cpu().numpy(), gld[key + "_t"])
        sg = int(gld[key + "_sign"])
        assert f.isPositive() == (sg in (0, 2)) and f.isNegative() == (sg in (1, 2))
        ldlt_check_factor(gld, key, f.matrixLDLT().cpu().numpy(),
                          1e-10 if kind == "sym" else 1e-13)
        x = f.solve(torch.from_numpy(b).cuda()).cpu().numpy()
        xr = gld[key + "_x"]
        assert np.abs(x - xr).max() <= 1e-8 * max(1.0, np.abs(xr).max())
        B = np.stack([b, 2 * b, -b], axis=1)
        X = f.solve(torch.from_numpy(np.asfortranarray(B)).cuda()).cpu().numpy()
        np.testing.assert_allclose(X, np.stack([x, 2 * x, -x], axis=1), rtol=1e-12,
                                   atol=1e-12 * np.abs(x).max())
        r = f.reconstructedMatrix().cpu().numpy()
        assert np.abs(r - a).max() <= 1e-12 * np.abs(a).max() * 24


def test_ldlt_graph_replays():
    """Factor + solve captured in one graph (no host synchronisation in
    either entry point): three replays, with the inputs restored and eager
    work in between, each equal to the eager result bit for bit."""
    from eigen_amd._lib import check, lib
    from eigen_amd.dense import _p, handle
    n, offA, A0, offb, b0 = inputs.packed_spd_gpu(5000, [6, 9, 12, 15, 18, 24], seed=13)
    A0, b0 = A0.to(torch.float64), b0.to(torch.float64)
    offT = offb.to(torch.int64)
    count = n.numel()
    # eager reference
    Ae, be = A0.clone(), b0.clone()
    Te, _, se, ie = batched.ldlt_batched(offA, n, Ae, offT, max_n=24)
    batched.ldlt_solve_batched(offA, n, Ae, offT, Te, offT, be, max_n=24)
    torch.cuda.synchronize()
    assert (ie == 0).all() and (se == 0).all()
    # captured: every buffer allocated before the capture
    A, b = A0.clone(), b0.clone()
    T = torch.zeros_like(Te)
    sign = torch.empty_like(se)
    info = torch.empty_like(ie)

    def step():
        h = handle()
        check("eamd_dldlt_batched",
              lib().eamd_dldlt_batched(h, count, _p(offA), _p(n), 24, _p(A), _p(offT), _p(T),
                                       _p(sign), _p(info)))
        check("eamd_dldlt_solve_batched",
              lib().eamd_dldlt_solve_batched(h, count, _p(offA), _p(n), 24, _p(A), _p(offT),
                                             _p(T), _p(offT), _p(b)))

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()  # bind the handle to this stream outside the capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        step()
    for rep in range(3):
        A.copy_(A0)
        b.copy_(b0)
        info.fill_(7)
        junk = torch.randn(1000, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(A, Ae) and torch.equal(T, Te), f"replay {rep}"
        assert torch.equal(b, be) and torch.equal(info, ie) and torch.equal(sign, se)
        del junk


def test_ldlt_rank_update_batched_against_reference_fixture():
    """LDLT::rankUpdate (LDLT.h:403-449, 529-555) of the batched GPU factors
    of every ldltrank.npz case in one launch: the updated factor within
    1e-10 (indefinite) / 1e-13 (definite) of Eigen's, its strict upper part
    and transpositions untouched, and the oracle's update of the same GPU
    factor matched to 1e-12; then LDLT.rankUpdate on a 200 x 200 factor
    (beyond the batched factorization's 64, within the update's 256) made
    by the oracle, against the oracle's update."""
    import os
    from tests.test_oracle import LDLT_RANK_SIGMAS, LDLT_RANK_SIZES
    gr = np.load(os.path.join(os.path.dirname(__file__), "golden", "ldltrank.npz"),
                 allow_pickle=False)
    cases = [(kind, k, tag) for kind in LDLT_KINDS for k in LDLT_RANK_SIZES
             for tag in sorted(LDLT_RANK_SIGMAS)]
    mats = [inputs.ldlt_matrix(kind, k) for kind, k, _ in cases]
    ws = [inputs.normal(9000 + k, k) for _, k, _ in cases]
    offA, n, A, T, offT, sign, info, off, sizes = run_ldlt(mats)
    f0 = [unpack(A, off, sizes, i) for i in range(len(cases))]
    Th0 = T.cpu().numpy().copy()
    ow = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    offw = torch.from_numpy(ow).cuda()
    W = torch.from_numpy(np.concatenate(ws)).cuda()
    sig = torch.tensor([LDLT_RANK_SIGMAS[tag] for _, _, tag in cases], **D64)
    print("ldlt rank update: batch of", len(cases), flush=True)
    rinfo = batched.ldlt_rank_update_batched(offA, n, A, offT, T, offw, W, sig)
    assert (rinfo.cpu().numpy() == 0).all()
    np.testing.assert_array_equal(T.cpu().numpy(), Th0)
    np.testing.assert_array_equal(W.cpu().numpy(), np.concatenate(ws))
    ot = offT.cpu().numpy()
    for i, (kind, k, tag) in enumerate(cases):
        key = f"{kind}{k}{tag}"
        f = unpack(A, off, sizes, i)
        ldlt_check_factor(gr, key, f, 1e-10 if kind == "sym" else 1e-13)
        np.testing.assert_array_equal(np.triu(f, 1), np.triu(mats[i], 1))
        ref = po.ldlt_rank_update(f0[i], Th0[ot[i]:ot[i] + k], ws[i], LDLT_RANK_SIGMAS[tag])
        scale = max(1.0, np.abs(np.tril(ref)).max())
        assert np.abs(np.tril(f) - np.tril(ref)).max() <= 1e-12 * scale, key
    k = 200
    a = inputs.ldlt_matrix("pos", k)
    fo, to, _, ok = po.ldlt(a)
    assert ok
    w = inputs.normal(9200, k)
    F = ea.LDLT()
    F._m = ea.colmajor(torch.from_numpy(np.asfortranarray(fo)))
    F._t = torch.from_numpy(np.ascontiguousarray(to)).cuda()
    F._sign, F._info = 0, ea.Success
    print("ldlt rank update: n = 200", flush=True)
    F.rankUpdate(torch.from_numpy(w), -0.5)
    ref = po.ldlt_rank_update(fo, to, w, -0.5)
    got = F.matrixLDLT().cpu().numpy()
    assert np.abs(np.tril(got) - np.tril(ref)).max() <= 1e-12 * max(1.0, np.abs(ref).max())
