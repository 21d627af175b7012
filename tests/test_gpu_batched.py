"""GPU tests of the batched small SPD solves (LLT<MatrixXf> per system):
against the oracle's per-system restatement, against Eigen's own solutions
on the committed 1,000-system sample, the failing-column rule on non-SPD
systems, every supported size, and replays of the captured graph."""
import os

import numpy as np
import pytest
import torch

from eigen_amd import batched
from oracle import pyoracle as po
from tests import inputs

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(__file__), "golden")


def to_dev(*arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


def test_batched_llt_matches_reference_sample(gpu):
    g = np.load(os.path.join(G, "batched.npz"), allow_pickle=False)
    n, offA, A, offb, b = inputs.packed_systems(11, 1000, [6, 9, 12, 15, 18, 24], np.float32,
                                                True)
    dn, doa, dA, dob, db = to_dev(n, offA, A, offb, b)
    info = batched.spotrf_solve_batched(doa, dn, dA, dob, db)
    assert (info.cpu().numpy() == -1).all()
    np.testing.assert_allclose(db.cpu().numpy(), g["llt_x"], rtol=2e-4, atol=2e-5)


@pytest.mark.parametrize("sizes", [list(range(1, 33)), [1], [32], [7, 24]])
def test_batched_llt_all_sizes_against_oracle(gpu, sizes):
    n, offA, A, offb, b = inputs.packed_systems(21, 3000, sizes, np.float32, True)
    x_ref, info_ref = po.batched_llt_solve_f(offA, n, A, offb, b)
    dn, doa, dA, dob, db = to_dev(n, offA, A, offb, b)
    info = batched.spotrf_solve_batched(doa, dn, dA, dob, db)
    np.testing.assert_array_equal(info.cpu().numpy(), info_ref)
    np.testing.assert_allclose(db.cpu().numpy(), x_ref, rtol=2e-4, atol=2e-5)


def test_batched_llt_failing_columns(gpu):
    n, offA, A, offb, b = inputs.packed_systems(22, 500, [6, 9, 18, 24], np.float32, True)
    bad = inputs.rng(5).choice(500, 60, replace=False)
    for i in bad:
        k = int(n[i])
        c = int(inputs.rng(int(i)).integers(0, k))
        A[offA[i] + c + c * k] = -1000.0
    x_ref, info_ref = po.batched_llt_solve_f(offA, n, A, offb, b)
    dn, doa, dA, dob, db = to_dev(n, offA, A, offb, b)
    info = batched.spotrf_solve_batched(doa, dn, dA, dob, db).cpu().numpy()
    np.testing.assert_array_equal(info, info_ref)
    ok = info == -1
    # solutions of the SPD systems are unaffected by their failing neighbours
    xs = db.cpu().numpy()
    for i in np.nonzero(ok)[0][:100]:
        k = int(n[i])
        np.testing.assert_allclose(xs[offb[i]:offb[i] + k], x_ref[offb[i]:offb[i] + k],
                                   rtol=2e-4, atol=2e-5)


def test_batched_llt_rejects_oversize(gpu):
    n, offA, A, offb, b = inputs.packed_systems(23, 10, [8], np.float32, True)
    n[3] = 40
    dn, doa, dA, dob, db = to_dev(n, offA, A, offb, b)
    with pytest.raises(Exception):
        batched.spotrf_solve_batched(doa, dn, dA, dob, db)


def test_batched_llt_graph_replays(gpu):
    """Replay the captured execution several times with eager work and
    synchronisations in between; every replay must equal the eager result."""
    n, offA, A, offb, b0 = inputs.packed_spd_gpu(20000, [6, 9, 12, 15, 18, 24], seed=9)
    ref = b0.clone()
    plan = batched.BatchedPlan(n, 32)
    info = torch.empty(n.numel(), dtype=torch.int32, device="cuda")
    plan.solve_llt(offA, A, offb, ref, info)
    torch.cuda.synchronize()
    b = b0.clone()
    info2 = torch.empty_like(info)
    g = batched.capture_llt(plan, offA, A, offb, b, info2)
    for rep in range(3):
        b.copy_(b0)
        junk = torch.randn(1000, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(b, ref), f"replay {rep}"
        assert (info2 == -1).all()
        del junk
    for s in (6, 9, 12, 15, 18, 24):
        assert plan.lanes(s) in (1, 2, 4, 8, 16)
