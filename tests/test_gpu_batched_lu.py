"""GPU tests of the batched uneven LU solve (PartialPivLU<MatrixXd> per
system, computed as an FP32 LU with FP64 refinement): Eigen's own solutions
on the committed 48-system sample, the oracle's per-system solves at every
size 1..256, the normwise backward error bound of 1e-14 per right-hand side,
the FP64 refactor of systems too ill-conditioned for the FP32 factor,
singular systems, and replays of the captured planned execution."""
import os

import numpy as np
import pytest
import torch

from eigen_amd import batched
from oracle import pyoracle as po
from tests import inputs

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-14


def to_dev(*arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


def berr_per_rhs(n, offA, A, offB, nrhs, B, X):
    """max over systems and right-hand sides of ||b - Ax||_inf /
    (||A||_inf ||x||_inf + ||b||_inf), computed on the host in FP64."""
    worst = 0.0
    for i in range(len(n)):
        k = int(n[i])
        a = A[offA[i]:offA[i] + k * k].reshape(k, k, order="F")
        b = B[offB[i]:offB[i] + k * nrhs].reshape(k, nrhs, order="F")
        x = X[offB[i]:offB[i] + k * nrhs].reshape(k, nrhs, order="F")
        r = b - a @ x
        an = np.abs(a).sum(axis=1).max()
        for q in range(nrhs):
            den = an * np.abs(x[:, q]).max() + np.abs(b[:, q]).max()
            worst = max(worst, np.abs(r[:, q]).max() / den)
    return worst


def rel_diff(n, offB, nrhs, X, Xr):
    worst = 0.0
    for i in range(len(n)):
        k = int(n[i]) * nrhs
        x, xr = X[offB[i]:offB[i] + k], Xr[offB[i]:offB[i] + k]
        worst = max(worst, np.abs(x - xr).max() / np.abs(xr).max())
    return worst


def test_batched_lu_matches_reference_sample(gpu):
    g = np.load(os.path.join(G, "batched.npz"), allow_pickle=False)
    n, offA, A, offB, B = inputs.packed_systems(12, 48, (16, 256), np.float64, False, nrhs=4)
    assert inputs.sha256(A) == str(g["lu_sha"])
    dn, doa, dA, dob, dB = to_dev(n, offA, A, offB, B)
    X, info, iters = batched.dgesv_ir_batched(doa, dn, dA, dob, 4, dB)
    X = X.cpu().numpy()
    assert (info.cpu().numpy() <= 1).all()
    assert berr_per_rhs(n, offA, A, offB, 4, B, X) <= TOL
    assert rel_diff(n, offB, 4, X, g["lu_x"]) <= 1e-9


@pytest.mark.parametrize("sizes,nrhs,count", [((1, 256), 1, 700), ((1, 40), 8, 600),
                                              ((200, 256), 3, 64), ([16, 64, 65, 128, 129, 256], 4, 120)])
def test_batched_lu_against_oracle(gpu, sizes, nrhs, count):
    n, offA, A, offB, B = inputs.packed_systems(31, count, sizes, np.float64, False, nrhs=nrhs)
    Xr, info_r = po.batched_lu_solve_d(offA, n, A, offB, nrhs, B)
    dn, doa, dA, dob, dB = to_dev(n, offA, A, offB, B)
    X, info, iters = batched.dgesv_ir_batched(doa, dn, dA, dob, nrhs, dB)
    X = X.cpu().numpy()
    info = info.cpu().numpy()
    assert (info <= 1).all(), np.nonzero(info > 1)
    assert berr_per_rhs(n, offA, A, offB, nrhs, B, X) <= TOL
    assert rel_diff(n, offB, nrhs, X, Xr) <= 1e-8
    # the FP32 factor refines almost every well-conditioned random system
    assert (info == 0).mean() > 0.95
    assert iters.cpu().numpy().max() <= 10


def _ill_conditioned(k, kappa, seed):
    g = inputs.rng(seed)
    q1, _ = np.linalg.qr(g.standard_normal((k, k)))
    q2, _ = np.linalg.qr(g.standard_normal((k, k)))
    s = np.logspace(0, -np.log10(kappa), k)
    return (q1 * s) @ q2.T


def test_batched_lu_refactors_ill_conditioned(gpu):
    """kappa = 1e10 is beyond FP32 refinement: those systems must come back
    through the FP64 refactor (info 1) and still meet the bound."""
    n, offA, A, offB, B = inputs.packed_systems(33, 40, (20, 120), np.float64, False, nrhs=2)
    bad = [3, 17, 29]
    for i in bad:
        k = int(n[i])
        A[offA[i]:offA[i] + k * k] = _ill_conditioned(k, 1e10, i).ravel(order="F")
    dn, doa, dA, dob, dB = to_dev(n, offA, A, offB, B)
    X, info, _ = batched.dgesv_ir_batched(doa, dn, dA, dob, 2, dB)
    info = info.cpu().numpy()
    X = X.cpu().numpy()
    assert set(np.nonzero(info == 1)[0]) == set(bad)
    assert (info[[i for i in range(40) if i not in bad]] == 0).all()
    assert berr_per_rhs(n, offA, A, offB, 2, B, X) <= TOL


def test_batched_lu_singular_reports_not_converged(gpu):
    n, offA, A, offB, B = inputs.packed_systems(34, 20, (8, 70), np.float64, False, nrhs=1)
    for i in (2, 11):
        k = int(n[i])
        a = A[offA[i]:offA[i] + k * k].reshape(k, k, order="F").copy()
        a[:, 5] = 0.0  # an exactly zero column
        A[offA[i]:offA[i] + k * k] = a.ravel(order="F")
    dn, doa, dA, dob, dB = to_dev(n, offA, A, offB, B)
    X, info, _ = batched.dgesv_ir_batched(doa, dn, dA, dob, 1, dB)
    info = info.cpu().numpy()
    assert info[2] == 2 and info[11] == 2
    ok = [i for i in range(20) if i not in (2, 11)]
    assert (info[ok] == 0).all()


def test_batched_lu_rejects_bad_arguments(gpu):
    n, offA, A, offB, B = inputs.packed_systems(35, 10, (8, 20), np.float64, False, nrhs=1)
    n[4] = 300
    dn, doa, dA, dob, dB = to_dev(n, offA, A, offB, B)
    with pytest.raises(Exception):
        batched.dgesv_ir_batched(doa, dn, dA, dob, 1, dB)
    n[4] = 10
    dn, = to_dev(n)
    with pytest.raises(Exception):
        batched.dgesv_ir_batched(doa, dn, dA, dob, 9, dB)


def test_batched_lu_planned_graph_replays(gpu):
    """The planned execution captured in a graph; every replay, with eager
    work and synchronisations in between, equals the eager result."""
    n, offA, A, offB, B = inputs.packed_systems(36, 300, (16, 256), np.float64, False, nrhs=4)
    dn, doa, dA, dob, dB = to_dev(n, offA, A, offB, B)
    plan = batched.BatchedLuPlan(dn, 4)
    for k in (16, 100, 256):
        assert plan.threads(k) in (0, 64, 128, 256)
    X = torch.empty_like(dB)
    info = torch.empty(300, dtype=torch.int32, device="cuda")
    iters = torch.empty_like(info)
    plan.solve(doa, dn, dA, dob, dB, X, info, iters)
    torch.cuda.synchronize()
    ref = X.clone()
    assert berr_per_rhs(n, offA, A, offB, 4, B, ref.cpu().numpy()) <= TOL
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    X2 = torch.empty_like(X)
    info2 = torch.empty_like(info)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.solve(doa, dn, dA, dob, dB, X2, info2, None)
    for rep in range(3):
        X2.zero_()
        junk = torch.randn(1000, device="cuda").sum()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(X2, ref), f"replay {rep}"
        assert torch.equal(info2, info)
        del junk
