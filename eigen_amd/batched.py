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


class BatchedLuPlan:
    """Plan of a batched LU + FP64 refinement solve (sizes n: int32 device
    tensor, 1..256; nrhs right-hand sides per system)."""

    def __init__(self, n, nrhs):
        self.count = int(n.numel())
        self.nrhs = int(nrhs)
        self.ptr = C.c_void_p()
        check("eamd_batched_lu_plan_create",
              lib().eamd_batched_lu_plan_create(handle(), self.count, _p(n), self.nrhs,
                                                C.byref(self.ptr)))

    def threads(self, n):
        """Workgroup size the autotuner chose for systems of size n."""
        return lib().eamd_batched_plan_lanes(self.ptr, n)

    def solve(self, offA, n, A, offB, B, X, info, iters=None):
        check("eamd_dgesv_ir_batched_planned",
              lib().eamd_dgesv_ir_batched_planned(
                  handle(), self.ptr, _p(offA), _p(n), _p(A), _p(offB), _p(B), _p(X), _p(info),
                  _p(iters) if iters is not None else None))

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


def llt_rank_update_batched(offL, n, L, offv, v, sigma=None, upper=False):
    """LLT::rankUpdate (LLT.h:459-472) for packed factors: L (FP64, factor i
    column-major n[i] x n[i] at offL[i]) updated in place so that
    L_i L_i^T = A_i + sigma_i v_i v_i^T; returns info (-1, the failing column
    of a downdate, or -2 for n outside 0..256)."""
    info = torch.empty(n.numel(), dtype=torch.int32, device=n.device)
    check("eamd_dllt_rank_update_batched",
          lib().eamd_dllt_rank_update_batched(handle(), n.numel(), _p(offL), _p(n), _p(L),
                                              _p(offv), _p(v),
                                              _p(sigma) if sigma is not None else None,
                                              1 if upper else 0, _p(info)))
    return info


def ldlt_batched(offA, n, A, offT=None, max_n=None):
    """LDLT<MatrixXd, Lower>::compute (LDLT.h:491-522) for packed symmetric
    matrices: A (FP64, matrix i column-major n[i] x n[i] at offA[i]) has its
    lower triangle replaced by matrixLDLT()'s.  Returns (T, sign, info):
    transpositions (int32, system i's n[i] entries at offT[i]; offT defaults
    to the prefix sums of n), sign per system (0 PositiveSemiDef,
    1 NegativeSemiDef, 2 ZeroSign, 3 Indefinite) and info (0 Success,
    1 NumericalIssue, -2 for n outside 0..max_n).  max_n (default: the
    largest n, read once on the host) is at most 64."""
    dev = n.device
    count = n.numel()
    if max_n is None:
        max_n = int(n.max()) if count else 0
    if offT is None:
        offT = (torch.cumsum(n.to(torch.int64), 0) - n.to(torch.int64)).contiguous()
    # the furthest end of any system's slots (offT need not be ascending)
    total = int((offT + n.to(torch.int64).clamp(min=0)).max()) if count else 0
    T = torch.zeros(max(total, 1), dtype=torch.int32, device=dev)
    sign = torch.empty(count, dtype=torch.int32, device=dev)
    info = torch.empty(count, dtype=torch.int32, device=dev)
    check("eamd_dldlt_batched",
          lib().eamd_dldlt_batched(handle(), count, _p(offA), _p(n), int(max_n), _p(A), _p(offT),
                                   _p(T), _p(sign), _p(info)))
    return T, offT, sign, info


def ldlt_solve_batched(offA, n, A, offT, T, offb, b, max_n=None):
    """LDLT::solve (LDLT.h:559-593) with ldlt_batched's factors: b (FP64,
    system i's right-hand side at offb[i]) overwritten with the solution;
    max_n as for ldlt_batched."""
    if max_n is None:
        max_n = int(n.max()) if n.numel() else 0
    check("eamd_dldlt_solve_batched",
          lib().eamd_dldlt_solve_batched(handle(), n.numel(), _p(offA), _p(n), int(max_n), _p(A),
                                         _p(offT), _p(T), _p(offb), _p(b)))
    return b


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


def packed_general_gpu(count, lo, hi, nrhs, seed, device="cuda"):
    """Packed FP64 general systems, A_i iid N(0,1) with n_i uniform integer in
    [lo, hi], B_i n_i x nrhs N(0,1).  Returns n, offA, A, offB, B (device)."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    n = torch.randint(lo, hi + 1, (count,), generator=g, device=device, dtype=torch.int64)
    offA = torch.zeros(count, dtype=torch.int64, device=device)
    offA[1:] = torch.cumsum(n * n, 0)[:-1]
    offB = torch.zeros(count, dtype=torch.int64, device=device)
    offB[1:] = torch.cumsum(n * nrhs, 0)[:-1]
    A = torch.randn(int((n * n).sum()), generator=g, device=device, dtype=torch.float64)
    B = torch.randn(int((n * nrhs).sum()), generator=g, device=device, dtype=torch.float64)
    return n.to(torch.int32), offA, A, offB, B


def _sample_berr(n, offA, A, offB, nrhs, B, X, idx):
    """max normwise backward error over the sampled systems (FP64, host)."""
    import numpy as np
    nn, oa, ob = n.cpu().numpy(), offA.cpu().numpy(), offB.cpu().numpy()
    worst = 0.0
    for i in idx:
        k = int(nn[i])
        a = A[int(oa[i]):int(oa[i]) + k * k].cpu().numpy().reshape(k, k, order="F")
        b = B[int(ob[i]):int(ob[i]) + k * nrhs].cpu().numpy().reshape(k, nrhs, order="F")
        x = X[int(ob[i]):int(ob[i]) + k * nrhs].cpu().numpy().reshape(k, nrhs, order="F")
        a, b, x = a.astype(np.float64), b.astype(np.float64), x.astype(np.float64)
        r = b - a @ x
        an = np.abs(a).sum(axis=1).max()
        for q in range(nrhs):
            den = an * np.abs(x[:, q]).max() + np.abs(b[:, q]).max()
            worst = max(worst, float(np.abs(r[:, q]).max() / den))
    return worst


def _loop_rate(n, offA, A, offB, nrhs, B, idx, kind):
    """Per-matrix launch loop of the plain (non-batched) GPU path over the
    sampled systems: PartialPivLU(A_i).solve(B_i) or LLT(A_i).solve(b_i) in
    FP64 through the drop-in entry points; returns solves per second."""
    import time

    from . import dense
    nn, oa, ob = n.cpu().numpy(), offA.cpu().numpy(), offB.cpu().numpy()
    mats = []
    for i in idx:
        k = int(nn[i])
        a = A[int(oa[i]):int(oa[i]) + k * k].view(k, k).t().to(torch.float64)
        b = B[int(ob[i]):int(ob[i]) + k * nrhs].view(nrhs, k).t().to(torch.float64)
        mats.append((a, b))
    for a, b in mats[:5]:  # warm the entry points
        (dense.PartialPivLU(a) if kind == "lu" else dense.LLT(a)).solve(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in mats:
        (dense.PartialPivLU(a) if kind == "lu" else dense.LLT(a)).solve(b)
    torch.cuda.synchronize()
    return len(mats) / (time.perf_counter() - t0)


def bench_lu_fields(count=65536, nrhs=4, steps=3):
    """Fields of the benchmark line for the batched uneven LU + FP64
    refinement workload (n_i uniform in [16, 256], 4 right-hand sides)."""
    import time

    import numpy as np
    n, offA, A, offB, B = packed_general_gpu(count, 16, 256, nrhs, seed=5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = BatchedLuPlan(n, nrhs)
    torch.cuda.synchronize()
    plan_ms = (time.perf_counter() - t0) * 1e3
    X = torch.empty_like(B)
    info = torch.empty(count, dtype=torch.int32, device="cuda")
    iters = torch.empty_like(info)
    plan.solve(offA, n, A, offB, B, X, info, iters)  # warm-up
    torch.cuda.synchronize()
    times = []
    for _ in range(steps):
        t1 = time.perf_counter()
        plan.solve(offA, n, A, offB, B, X, info, iters)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t1)
    dt = min(times)
    n64 = n.to(torch.int64).double()
    flops = float((2.0 / 3.0 * n64 ** 3 + 2.0 * nrhs * n64 ** 2).sum())
    bytes_moved = float((8.0 * n64 ** 2 + 16.0 * nrhs * n64 + 24.0).sum())
    idx = np.random.Generator(np.random.Philox(6)).choice(count, 200, replace=False)
    loop_idx = idx[:100]
    loop = _loop_rate(n, offA, A, offB, nrhs, B, loop_idx, "lu")
    inf = info.cpu().numpy()
    return {
        "workload": f"{count} general systems, n uniform integer in [16, 256], FP64 N(0,1), "
                    f"{nrhs} RHS, FP32 LU on MFMA + FP64 refinement to 1e-14",
        "solves_per_s": round(count / dt, 1),
        "ms": round(dt * 1e3, 3),
        "plan_ms": round(plan_ms, 2),
        "tflops_fp64_equiv": round(flops / dt / 1e12, 3),
        "hbm_gbs": round(bytes_moved / dt / 1e9, 1),
        "hbm_frac": round(bytes_moved / dt / 8e12, 4),
        "sample_backward_error_max": _sample_berr(n, offA, A, offB, nrhs, B, X, idx),
        "refined_from_fp32": int((inf == 0).sum()),
        "fp64_refactor": int((inf == 1).sum()),
        "not_converged": int((inf == 2).sum()),
        "max_refinement_steps": int(iters.max().item()),
        "per_matrix_loop_solves_per_s": round(loop, 1),
        "speedup_vs_per_matrix_loop": round(count / dt / loop, 1),
        "threads_per_system": {k: plan.threads(k) for k in (16, 64, 128, 192, 256)},
    }


def bench_rank_update_fields(n, offA, A, offv, steps=5):
    """Rank-one updates (LLT::rankUpdate, sigma = 1) of the batched workload's
    factors in FP64: the packed SPD matrices' lower triangles stand in for
    factors (a positive diagonal is all an update with sigma > 0 needs)."""
    import time
    count = n.numel()
    L = A.to(torch.float64)
    V = torch.randn(int(n.sum().item()), dtype=torch.float64, device=A.device,
                    generator=torch.Generator(device=A.device).manual_seed(5))
    offv = offv.to(torch.int64)
    times = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        info = llt_rank_update_batched(offA, n, L, offv, V)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t1)
    dt = min(times)
    nn = n.to(torch.int64)
    bytes_moved = float((8 * (nn * (nn + 1)) + 8 * nn).sum().item())  # L read + write, v
    return {"workload": f"{count} FP64 factors of the batched LLT sizes, one rank-one "
                        "update each (sigma = 1), packed + offsets",
            "updates_per_s": round(count / dt, 1), "ms": round(dt * 1e3, 3),
            "hbm_gbs": round(bytes_moved / dt / 1e9, 1),
            "hbm_frac": round(bytes_moved / dt / 8e12, 4),
            "all_info_ok": bool((info == -1).all().item())}


def bench_ldlt_fields(n, offA, A, offb, b, steps=5):
    """LDLT compute + solve (LDLT.h:491-522, 559-593) of the batched
    workload's SPD matrices in FP64, one right-hand side each; the packed
    inputs are restored (untimed) before every timed step."""
    import time
    count = n.numel()
    A0 = A.to(torch.float64)
    b0 = b.to(torch.float64)
    offT = offb.to(torch.int64)
    max_n = int(n.max().item())
    Aw, bw = A0.clone(), b0.clone()
    times = []
    for _ in range(steps):
        Aw.copy_(A0)
        bw.copy_(b0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        T, _, sign, info = ldlt_batched(offA, n, Aw, offT, max_n=max_n)
        ldlt_solve_batched(offA, n, Aw, offT, T, offT, bw, max_n=max_n)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t1)
    dt = min(times)
    return {"workload": f"{count} FP64 symmetric systems of the batched LLT sizes, LDLT "
                        "compute + solve (one right-hand side), packed + offsets",
            "systems_per_s": round(count / dt, 1), "ms": round(dt * 1e3, 3),
            "all_info_ok": bool((info == 0).all().item()),
            "all_positive": bool((sign == 0).all().item())}


def bench_fields(count=1_000_000, steps=5):
    """Extra fields of the benchmark line for the batched workloads."""
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
    try:
        loop = _loop_rate(n, offA, A, offb, 1, b0, idx[:300], "llt")
        out["batched_llt"]["per_matrix_loop_solves_per_s"] = round(loop, 1)
        out["batched_llt"]["speedup_vs_per_matrix_loop"] = round(count / dt / loop, 1)
    except Exception as e:  # reported, never hides the batched result
        out["batched_llt"]["per_matrix_loop_error"] = str(e)[:200]
    del b, x
    try:
        out["batched_llt_rank_update"] = bench_rank_update_fields(n, offA, A, offb, steps)
    except Exception as e:
        out["batched_llt_rank_update"] = {"error": str(e)[:300]}
    try:
        out["batched_ldlt"] = bench_ldlt_fields(n, offA, A, offb, b0, steps)
    except Exception as e:  # reported, not fatal to the headline line
        out["batched_ldlt"] = {"error": str(e)[:300]}
    del b0
    del A
    try:
        out["batched_lu"] = bench_lu_fields()
    except Exception as e:
        out["batched_lu"] = {"error": str(e)[:300]}
    return out
