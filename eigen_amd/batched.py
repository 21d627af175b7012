"""Batched small solves over packed variable-size systems (device tensors).

Per system this is ``LLT<MatrixXf>(A_i).solve(b_i)`` (LLT) or
``PartialPivLU<MatrixXd>(A_i).solve(B_i)`` (LU with FP64 refinement); the
systems are packed column-major back to back, located by an offset array,
and solved in one captured launch sequence.
"""
import ctypes as C

import torch

from ._lib import check, lib
from .dense import handle


def _p(t):
    return C.c_void_p(t.data_ptr())


class BatchedPlan:
    """Size-bucketing plan for a packed batch (sizes n: int32 device tensor)."""

    def __init__(self, n, max_n=32):
        self.count = int(n.numel())
        self.ptr = C.c_void_p()
        check("eamd_batched_plan_create",
              lib().eamd_batched_plan_create(handle(), self.count, _p(n), max_n,
                                             C.byref(self.ptr)))

    def lanes(self, n):
        return lib().eamd_batched_plan_lanes(self.ptr, n)

    def solve_llt(self, offA, A, offb, b, info):
        """b <- A_i^{-1} b_i per system (in place); info[i] = failing column or -1."""
        check("eamd_spotrf_solve_batched_planned",
              lib().eamd_spotrf_solve_batched_planned(handle(), self.ptr, _p(offA), _p(A),
                                                      _p(offb), _p(b), _p(info)))

    def __del__(self):
        try:
            if self.ptr:
                lib().eamd_batched_plan_destroy(self.ptr)
        except Exception:
            pass


def spotrf_solve_batched(offA, n, A, offb, b):
    """One-shot batched LLT solve: b overwritten with x; returns info."""
    info = torch.empty(n.numel(), dtype=torch.int32, device=n.device)
    check("eamd_spotrf_solve_batched",
          lib().eamd_spotrf_solve_batched(handle(), n.numel(), _p(offA), _p(n), _p(A), _p(offb),
                                          _p(b), _p(info)))
    return info


def dgesv_ir_batched(offA, n, A, offB, nrhs, B):
    """Batched FP32-LU + FP64-refinement solve; returns (X, info, iters)."""
    X = torch.empty_like(B)
    info = torch.empty(n.numel(), dtype=torch.int32, device=n.device)
    iters = torch.empty(n.numel(), dtype=torch.int32, device=n.device)
    check("eamd_dgesv_ir_batched",
          lib().eamd_dgesv_ir_batched(handle(), n.numel(), _p(offA), _p(n), _p(A), _p(offB),
                                      nrhs, _p(B), _p(X), _p(info), _p(iters)))
    return X, info, iters


def capture_llt(plan, offA, A, offb, b, info):
    """A torch CUDA graph replaying the planned batched LLT on these buffers."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.solve_llt(offA, A, offb, b, info)  # warm the kernels outside capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    # capture on the stream the library handle is already bound to, so binding
    # inside the capture is a no-op (no synchronisation while capturing)
    with torch.cuda.graph(g, stream=s):
        plan.solve_llt(offA, A, offb, b, info)
    return g


# ------------------------------------------------------------- generators --
def packed_spd_gpu(count, sizes, seed, device="cuda"):
    """Packed FP32 SPD systems A_i = B B^T + n_i I (B iid N(0,1)), b_i N(0,1);
    sizes drawn uniformly from `sizes`.  Returns n, offA, A, offb, b (device)."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    sz = torch.tensor(sizes, dtype=torch.int32, device=device)
    pick = torch.randint(0, len(sizes), (count,), generator=g, device=device)
    n = sz[pick]
    n64 = n.to(torch.int64)
    offA = torch.zeros(count, dtype=torch.int64, device=device)
    offA[1:] = torch.cumsum(n64 * n64, 0)[:-1]
    offb = torch.zeros(count, dtype=torch.int64, device=device)
    offb[1:] = torch.cumsum(n64, 0)[:-1]
    A = torch.empty(int((n64 * n64).sum()), dtype=torch.float32, device=device)
    for s in sizes:
        idx = torch.nonzero(n == s).flatten()
        k = idx.numel()
        if k == 0:
            continue
        for c0 in range(0, k, 65536):
            ii = idx[c0:c0 + 65536]
            Bm = torch.randn(ii.numel(), s, s, generator=g, device=device, dtype=torch.float32)
            Am = torch.bmm(Bm, Bm.transpose(1, 2).contiguous())
            Am += s * torch.eye(s, device=device, dtype=torch.float32)
            flat = offA[ii][:, None] + torch.arange(s * s, device=device)[None, :]
            A[flat.flatten()] = Am.transpose(1, 2).reshape(-1)
    b = torch.randn(int(n64.sum()), generator=g, device=device, dtype=torch.float32)
    return n, offA, A, offb, b


def bench_fields(count=1_000_000, steps=5):
    """Extra fields of the benchmark line for the batched LLT workload."""
    import time

    import numpy as np
    out = {}
    sizes = [6, 9, 12, 15, 18, 24]
    n, offA, A, offb, b0 = packed_spd_gpu(count, sizes, seed=4)
    b = b0.clone()
    info = torch.empty(count, dtype=torch.int32, device="cuda")
    t0 = time.perf_counter()
    plan = BatchedPlan(n, 32)
    torch.cuda.synchronize()
    plan_ms = (time.perf_counter() - t0) * 1e3
    g = capture_llt(plan, offA, A, offb, b, info)
    # each replay solves in place: restore b before each timed replay
    times = []
    for _ in range(steps):
        b.copy_(b0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g.replay()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t1)
    dt = min(times)
    x = b
    # accuracy on a 1,000-system sample: FP64 residual of the FP32 solution
    idx = np.random.Generator(np.random.Philox(3)).choice(count, 1000, replace=False)
    nn = n.cpu().numpy()
    oa = offA.cpu().numpy()
    ob = offb.cpu().numpy()
    Ah = A.cpu().numpy()
    bh = b0.cpu().numpy()
    xh = x.cpu().numpy()
    worst = 0.0
    for i in idx:
        k = int(nn[i])
        a = Ah[oa[i]:oa[i] + k * k].reshape(k, k, order="F").astype(np.float64)
        bb = bh[ob[i]:ob[i] + k].astype(np.float64)
        xx = xh[ob[i]:ob[i] + k].astype(np.float64)
        r = np.abs(a @ xx - bb).max()
        den = np.abs(a).sum(axis=1).max() * np.abs(xx).max() + np.abs(bb).max()
        worst = max(worst, r / den)
    bytes_moved = float(4 * (nn.astype(np.int64) ** 2 + 2 * nn).sum() + 8 * count)
    out["batched_llt"] = {
        "workload": f"{count} SPD systems, n uniform in {sizes}, FP32, packed + offsets, "
                    "1 RHS, one captured graph",
        "solves_per_s": round(count / dt, 1),
        "ms": round(dt * 1e3, 4),
        "plan_ms": round(plan_ms, 2),
        "hbm_gbs": round(bytes_moved / dt / 1e9, 1),
        "hbm_frac": round(bytes_moved / dt / 8e12, 4),
        "lanes_per_system": {s: plan.lanes(s) for s in sizes},
        "sample_backward_error_max": worst,
        "all_info_ok": bool((info == -1).all().item()),
    }
    return out
