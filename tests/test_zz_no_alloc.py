"""GPU test of eamd_spotrf_solve_batched_noalloc (the allocation-free
variant of the batched LLT solve) through the C ABI: two calls in a row,
the second with fewer systems, each equal to eamd_spotrf_solve_batched on
the same inputs; sizes beyond its maximum are refused."""
import ctypes as C

import numpy as np
import pytest
import torch

import eigen_amd as ea
from eigen_amd import batched
from tests import inputs

pytestmark = pytest.mark.gpu


def _p(t):
    return C.c_void_p(t.data_ptr())


def run_noalloc(n, offA, A, offb, b):
    info = torch.full((n.numel(),), 7, dtype=torch.int32, device="cuda")
    st = ea.lib().eamd_spotrf_solve_batched_noalloc(ea.handle(), n.numel(), _p(offA), _p(n),
                                                    _p(A), _p(offb), _p(b), _p(info))
    assert st == 0, ea.lib().eamd_last_error()
    torch.cuda.synchronize()
    return info


@pytest.mark.parametrize("count,seed", [(30000, 31), (2500, 32)])
def test_noalloc_equals_planned_entry(gpu, count, seed):
    sizes = list(range(1, 33))
    n, offA, A, offb, b0 = inputs.packed_spd_gpu(count, sizes, seed=seed)
    # a few failing systems too: info must agree, and their b stay as they were
    nh = n.cpu().numpy()
    oa = offA.cpu().numpy()
    for i in range(0, count, 997):
        k = int(nh[i])
        A[int(oa[i]) + (k - 1) * (k + 1)] = -1.0
    b_ref = b0.clone()
    info_ref = batched.spotrf_solve_batched(offA, n, A, offb, b_ref)
    torch.cuda.synchronize()
    b = b0.clone()
    info = run_noalloc(n, offA, A, offb, b)
    np.testing.assert_array_equal(info.cpu().numpy(), info_ref.cpu().numpy())
    np.testing.assert_array_equal(b.cpu().numpy(), b_ref.cpu().numpy())
    assert (info_ref.cpu().numpy() >= 0).any() and (info_ref.cpu().numpy() == -1).any()


def test_noalloc_refuses_beyond_maximum(gpu):
    n = torch.full((1, ), 4, dtype=torch.int32, device="cuda")
    offA = torch.zeros(1, dtype=torch.int64, device="cuda")
    A = torch.eye(4, dtype=torch.float32, device="cuda").reshape(-1)
    b = torch.ones(4, dtype=torch.float32, device="cuda")
    info = torch.empty(1, dtype=torch.int32, device="cuda")
    st = ea.lib().eamd_spotrf_solve_batched_noalloc(ea.handle(), (1 << 20) + 1, _p(offA),
                                                    _p(n), _p(A), _p(offA), _p(b), _p(info))
    assert st == -1002 and b"maximum" in ea.lib().eamd_last_error()
