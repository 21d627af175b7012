#!/usr/bin/env python3
"""Benchmark of the MI355X dense factor-and-solve path.

Headline (one JSON line on stdout): mixed-precision LU with iterative
refinement at n = 16384 (A iid N(0,1), one right-hand side), TFLOP/s by the
HPL count 2/3 n^3 + 2 n^2 over the wall time of factor + refinement to an
FP64 normwise backward error <= 1e-14.  A "step" is one such solve; inputs
are resident in HBM before the timed region.  The plain FP64 GPU LU
(factor + solve) is timed beside it for the speedup, the ill-conditioned
Q1 diag(s) Q2 matrix (s log-spaced 1..1e-6) is solved once for its report,
and the batched workloads are reported as extra fields when built.

roofline: the dominant kernel of the timed region (the trailing update of
the LU) -- algorithmic bytes (FP32 C read + write, packed FP16 operands read
once) divided by its device time, both from hipEvent pairs the library
records around its own launches on its stream.

cpu_baseline: the reference itself (Eigen 3.3.90 PartialPivLU<MatrixXd>,
compiled by oracle/Makefile into oracle/_ref) on the host cores, on a
bounded sample of the same workload (one n = 4096 FP64 factor + solve).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "n=16384 solve TFLOP/s mixed-IR vs FP64 @1e-14 bwd err; batched solves/s"
PEAK_HBM_GBS = 8000.0
PEAK_FP16_TFLOPS = 2500.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def hpl_flops(n):
    return 2.0 / 3.0 * n ** 3 + 2.0 * n ** 2


def householder_matrix_gpu(n, nref, seed, smin=1e-6):
    """A = Q1 diag(s) Q2, each Q a product of nref reflectors I - 2 v v^T / |v|^2."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    s = 10.0 ** (-torch.log10(torch.tensor(1.0 / smin, dtype=torch.float64)) *
                 torch.arange(n, dtype=torch.float64) / (n - 1))
    M = torch.diag(s).cuda()
    for side in (0, 1):
        for _ in range(nref):
            v = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
            v /= v.norm()
            if side == 0:
                w = v @ M
                M.addr_(v, w, alpha=-2.0)
            else:
                w = M @ v
                M.addr_(w, v, alpha=-2.0)
    return M.t().contiguous().t()  # column-major view


def profile_phase(ea, phase):
    import ctypes as C
    ms, by, fl, cnt = C.c_double(), C.c_double(), C.c_double(), C.c_int64()
    ea.lib().eamd_profile_read(ea.handle(), phase, C.byref(ms), C.byref(by), C.byref(fl),
                               C.byref(cnt))
    return ms.value, by.value, fl.value, cnt.value


def cpu_baseline(n=4096):
    """Eigen's PartialPivLU<MatrixXd> (factor + 1-RHS solve) on all host cores."""
    import numpy as np
    from oracle import pyoracle as po
    ref = po.ref()
    threads = len(os.sched_getaffinity(0))
    rng = np.random.Generator(np.random.Philox(7))
    a = np.asfortranarray(rng.standard_normal((n, n)))
    b = np.asfortranarray(rng.standard_normal((n, 1)))
    x = np.zeros((n, 1), order="F")
    lu = np.zeros((n, n), order="F")
    perm = np.zeros(n, dtype=np.int32)
    import ctypes as C
    det = C.c_double()
    if ref is not None:
        ref.eref_set_threads(threads)
        t0 = time.perf_counter()
        ref.eref_partialpivlu_d(n, a, n, 1, b, n, lu, perm, x, n, C.byref(det))
        dt = time.perf_counter() - t0
        kind = "reference"
        cores = ref.eref_max_threads()
    else:  # the restatement (single thread) when the reference was not built
        t0 = time.perf_counter()
        lu_, t, _, _ = po.lu(a)
        po.lu_solve(lu_, t, b)
        dt = time.perf_counter() - t0
        kind, cores = "port", 1
    out = {"value": round(hpl_flops(n) / dt / 1e12, 5), "unit": "TFLOP/s", "cores": cores,
           "kind": kind,
           "sample": f"PartialPivLU<MatrixXd> factor+solve, n={n} FP64 N(0,1), 1 RHS, "
                     f"{dt:.2f} s (HPL flop count)",
           "cpu_model": cpu_model()}
    if ref is not None:
        out.update(cpu_batched_legs(ref, cores))
    return out


def cpu_model():
    """The host CPU's model name and logical CPU count as the OS reports them."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name} ({os.cpu_count()} logical CPUs, {len(os.sched_getaffinity(0))} usable)"


def cpu_batched_legs(ref, cores):
    """The reference per system on the host cores (OpenMP over systems, Eigen
    serial inside: Parallelizer.h:123): LLT<MatrixXf> solves of the C4
    workload and PartialPivLU<MatrixXd> 4-RHS solves of a bounded sample of
    the C5 workload, on the same synthetic inputs as the GPU legs."""
    import numpy as np

    from tests import inputs
    legs = {}
    n, offA, A, offb, b = inputs.packed_spd_gpu(1_000_000, [6, 9, 12, 15, 18, 24], seed=4)
    nh, oah, Ah, obh = (n.cpu().numpy(), offA.cpu().numpy(), A.cpu().numpy(),
                        offb.cpu().numpy())
    bh = b.cpu().numpy()
    del n, offA, A, offb, b
    info = np.zeros(len(nh), dtype=np.int32)
    t0 = time.perf_counter()
    ref.eref_batched_llt_solve_f(len(nh), oah, nh, Ah, obh, bh, info)
    dt = time.perf_counter() - t0
    legs["batched_llt"] = {"solves_per_s": round(len(nh) / dt, 1), "cores": cores,
                           "sample": f"all {len(nh)} systems of the batched LLT workload, "
                                     f"{dt:.3f} s"}
    cnt = 8192
    n, offA, A, offB, B = inputs.packed_general_gpu(cnt, 16, 256, 4, seed=5)
    nh, oah, Ah, obh, Bh = (n.cpu().numpy(), offA.cpu().numpy(), A.cpu().numpy(),
                            offB.cpu().numpy(), B.cpu().numpy())
    info = np.zeros(cnt, dtype=np.int32)
    t0 = time.perf_counter()
    ref.eref_batched_lu_solve_d(cnt, oah, nh, Ah, obh, 4, Bh, info)
    dt = time.perf_counter() - t0
    legs["batched_lu"] = {"solves_per_s": round(cnt / dt, 1), "cores": cores,
                          "sample": f"{cnt} systems of the batched LU workload's distribution "
                                    f"(n uniform in [16, 256], 4 RHS), {dt:.3f} s"}
    return legs


def measured_traffic():
    """HBM bytes per launch of the dominant kernel, as recorded from the
    FETCH_SIZE / WRITE_SIZE counter runs (profiles/, each counter scaled by
    what it reads for a copy of known size); None when not recorded."""
    import re
    path = os.path.join(ROOT, "profiles", "mixed_update_kernel.n16384.traffic.md")
    try:
        with open(path) as f:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def _median_ms(fn, reps, warm=2):
    import statistics

    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return statistics.median(ts)


def dense_fields():
    """The plain FP64 drop-ins at their own configurations: PartialPivLU
    factor + solve at n = 1024 (pivots compared with the oracle's), GEMM
    8192^3, and LLT of A = G G^T / 8192 + I (n = 8192) with 16 RHS."""
    import numpy as np
    import torch

    import eigen_amd as ea
    from oracle import pyoracle as po
    out = {}
    g = torch.Generator(device="cuda")
    g.manual_seed(1024)
    n = 1024
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g).t()
    b = torch.randn(n, 1, dtype=torch.float64, device="cuda", generator=g)
    state = {}

    def lu_solve():
        state["lu"] = ea.PartialPivLU(A)
        state["x"] = state["lu"].solve(b)
    ms = _median_ms(lu_solve, 20, 3)
    _, t_ref, _, _ = po.lu(A.cpu().numpy())
    t = state["lu"].transpositionsP().cpu().numpy()
    out["lu_n1024"] = {"ms": round(ms, 3), "tflops": round(hpl_flops(n) / ms / 1e9, 4),
                       "backward_error": ea.backward_error(A, state["x"], b),
                       "pivots_differing_from_oracle": int((t != t_ref).sum())}
    del A, b, state
    n = 8192
    Ga = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g).t()
    Gb = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g).t()
    Cm = ea.colmajor(torch.empty(n, n, dtype=torch.float64, device="cuda"))
    ms = _median_ms(lambda: ea.gemm(Ga, Gb, Cm, beta=0.0), 5, 1)
    out["gemm_8192"] = {"ms": round(ms, 2), "tflops": round(2.0 * n ** 3 / ms / 1e9, 2),
                        "frac_fp64_peak": round(2.0 * n ** 3 / ms / 1e9 / 78.6, 4)}
    S = ea.gemm(Ga, Ga, transb="T", alpha=1.0 / n)
    S.diagonal().add_(1.0)
    B = torch.randn(n, 16, dtype=torch.float64, device="cuda", generator=g).t().contiguous().t()
    del Ga, Gb, Cm
    st = {}

    def llt():
        st["f"] = ea.LLT(S)
    ms_f = _median_ms(llt, 3, 1)
    ms_s = _median_ms(lambda: st.__setitem__("x", st["f"].solve(B)), 3, 1)
    out["llt_8192_16rhs"] = {"factor_ms": round(ms_f, 2),
                             "factor_tflops": round(n ** 3 / 3.0 / ms_f / 1e9, 2),
                             "solve_ms": round(ms_s, 2),
                             "backward_error": ea.backward_error(S, st["x"], B)}
    return out


# ------------------------------------------------------- batched workloads --
def _sample_berr(n, offA, A, offB, nrhs, B, X, idx):
    """max normwise backward error over the sampled systems (FP64, host)."""
    import numpy as np
    nn, oa, ob = n.cpu().numpy(), offA.cpu().numpy(), offB.cpu().numpy()
    Ah, Bh, Xh = A.cpu().numpy(), B.cpu().numpy(), X.cpu().numpy()
    worst = 0.0
    for i in idx:
        k = int(nn[i])
        a = Ah[int(oa[i]):int(oa[i]) + k * k].reshape(k, k, order="F").astype(np.float64)
        b = Bh[int(ob[i]):int(ob[i]) + k * nrhs].reshape(k, nrhs, order="F").astype(np.float64)
        x = Xh[int(ob[i]):int(ob[i]) + k * nrhs].reshape(k, nrhs, order="F").astype(np.float64)
        r = b - a @ x
        an = np.abs(a).sum(axis=1).max()
        for q in range(nrhs):
            den = an * np.abs(x[:, q]).max() + np.abs(b[:, q]).max()
            worst = max(worst, float(np.abs(r[:, q]).max() / den))
    return worst


def loop_rate_cabi(n, offA, A, offB, nrhs, B, idx, kind):
    """The per-matrix launch loop of the plain FP64 drop-in, through the C ABI
    (no Python objects in the loop): eamd_dgetrf + eamd_dgetrs ("lu") or
    eamd_dpotrf + eamd_dpotrs ("llt") on each sampled system, with device
    copies of the systems prepared before the timed loop; solves per second."""
    import ctypes as C

    import torch

    import eigen_amd as ea
    L, h = ea.lib(), ea.handle()
    nn, oa, ob = n.cpu().numpy(), offA.cpu().numpy(), offB.cpu().numpy()
    mats = []
    for i in idx:
        k = int(nn[i])
        a = A[int(oa[i]):int(oa[i]) + k * k].to(torch.float64).clone()
        b = B[int(ob[i]):int(ob[i]) + k * nrhs].to(torch.float64).clone()
        piv = torch.empty(k, dtype=torch.int32, device=A.device)
        mats.append((k, a, b, piv))
    nbt, fz, fail = C.c_int32(), C.c_int64(), C.c_int64()

    def run(batch):
        for k, a, b, piv in batch:
            if kind == "lu":
                L.eamd_dgetrf(h, k, k, a.data_ptr(), k, piv.data_ptr(), C.byref(nbt),
                              C.byref(fz))
                L.eamd_dgetrs(h, b"N", k, nrhs, a.data_ptr(), k, piv.data_ptr(), b.data_ptr(),
                              k)
            else:
                L.eamd_dpotrf(h, b"L", k, a.data_ptr(), k, C.byref(fail))
                L.eamd_dpotrs(h, b"L", k, nrhs, a.data_ptr(), k, b.data_ptr(), k)

    warm = [(k, a.clone(), b.clone(), p) for k, a, b, p in mats[:5]]
    run(warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(mats)
    torch.cuda.synchronize()
    return len(mats) / (time.perf_counter() - t0)


def batched_llt_fields(count=1_000_000, steps=5):
    """C4: 1M packed FP32 SPD systems, one persistent launch captured in a
    graph; returns (fields, inputs kept for the rank-update and LDLT legs)."""
    import numpy as np
    import torch

    from eigen_amd import batched
    from tests import inputs
    sizes = [6, 9, 12, 15, 18, 24]
    n, offA, A, offb, b0 = inputs.packed_spd_gpu(count, sizes, seed=4)
    b = b0.clone()
    info = torch.empty(count, dtype=torch.int32, device="cuda")
    t0 = time.perf_counter()
    plan = batched.BatchedPlan(n, 32)
    torch.cuda.synchronize()
    plan_ms = (time.perf_counter() - t0) * 1e3
    g = batched.capture_llt(plan, offA, A, offb, b, info)
    # each replay solves in place: restore b before each timed replay; the
    # replay is timed with events on the stream it runs on
    times = []
    for _ in range(steps):
        b.copy_(b0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) * 1e-3)
    dt = min(times)
    idx = np.random.Generator(np.random.Philox(3)).choice(count, 1000, replace=False)
    worst = _sample_berr(n, offA, A, offb, 1, b0, b, idx)
    nn = n.cpu().numpy().astype(np.int64)
    bytes_moved = float(4 * (nn ** 2 + 2 * nn).sum() + 8 * count)
    out = {
        "workload": f"{count} SPD systems, n uniform in {sizes}, FP32, packed + offsets, "
                    "1 RHS, one persistent launch captured in a graph",
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
        loop = loop_rate_cabi(n, offA, A, offb, 1, b0, idx[:300], "llt")
        out["per_matrix_loop_solves_per_s"] = round(loop, 1)
        out["speedup_vs_per_matrix_loop"] = round(count / dt / loop, 1)
    except Exception as e:  # reported, never hides the batched result
        out["per_matrix_loop_error"] = str(e)[:200]
    del b
    return out, (n, offA, A, offb, b0)


def batched_lu_fields(count=65536, nrhs=4, steps=3):
    """C5: batched uneven LU + FP64 refinement (n_i uniform in [16, 256],
    4 right-hand sides)."""
    import numpy as np
    import torch

    from eigen_amd import batched
    from tests import inputs
    n, offA, A, offB, B = inputs.packed_general_gpu(count, 16, 256, nrhs, seed=5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = batched.BatchedLuPlan(n, nrhs)
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
    loop = loop_rate_cabi(n, offA, A, offB, nrhs, B, idx[:100], "lu")
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


def rank_update_fields(n, offA, A, offv, steps=5):
    """Rank-one updates (LLT::rankUpdate, sigma = 1) of the C4 workload's
    factors in FP64: the packed SPD matrices' lower triangles stand in for
    factors (a positive diagonal is all an update with sigma > 0 needs)."""
    import torch

    from eigen_amd import batched
    count = n.numel()
    L = A.to(torch.float64)
    V = torch.randn(int(n.sum().item()), dtype=torch.float64, device=A.device,
                    generator=torch.Generator(device=A.device).manual_seed(5))
    offv = offv.to(torch.int64)
    times = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        info = batched.llt_rank_update_batched(offA, n, L, offv, V)
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


def ldlt_fields(n, offA, A, offb, b, steps=5):
    """LDLT compute + solve (LDLT.h:491-522, 559-593) of the C4 workload's SPD
    matrices in FP64, one right-hand side each; the packed inputs are
    restored (untimed) before every timed step."""
    import torch

    from eigen_amd import batched
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
        T, _, sign, info = batched.ldlt_batched(offA, n, Aw, offT, max_n=max_n)
        batched.ldlt_solve_batched(offA, n, Aw, offT, T, offT, bw, max_n=max_n)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t1)
    dt = min(times)
    return {"workload": f"{count} FP64 symmetric systems of the batched LLT sizes, LDLT "
                        "compute + solve (one right-hand side), packed + offsets",
            "systems_per_s": round(count / dt, 1), "ms": round(dt * 1e3, 3),
            "all_info_ok": bool((info == 0).all().item()),
            "all_positive": bool((sign == 0).all().item())}


def batched_fields():
    """Extra fields of the benchmark line for the batched workloads; each leg
    reports its own failure without hiding the others."""
    out = {}
    kept = None
    try:
        out["batched_llt"], kept = batched_llt_fields()
    except Exception as e:
        out["batched_llt"] = {"error": str(e)[:300]}
    if kept is not None:
        n, offA, A, offb, b0 = kept
        for key, fn in (("batched_llt_rank_update", lambda: rank_update_fields(n, offA, A, offb)),
                        ("batched_ldlt", lambda: ldlt_fields(n, offA, A, offb, b0))):
            try:
                out[key] = fn()
            except Exception as e:
                out[key] = {"error": str(e)[:300]}
        del n, offA, A, offb, b0, kept
    try:
        out["batched_lu"] = batched_lu_fields()
    except Exception as e:
        out["batched_lu"] = {"error": str(e)[:300]}
    return out


def fp64_fields(A, b, n, dt):
    """The plain FP64 GPU LU (factor + solve) on the headline input, and the
    mixed solver's speedup over it."""
    import torch

    from eigen_amd import mixed
    x64, rep64 = mixed.gesv_ir(A, b, precision="fp64")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    x64, rep64 = mixed.gesv_ir(A, b, precision="fp64")
    torch.cuda.synchronize()
    dt64 = time.perf_counter() - t1
    out = {"fp64_lu_tflops": round(hpl_flops(n) / dt64 / 1e12, 3),
           "fp64_lu_ms": round(dt64 * 1e3, 2),
           "fp64_backward_error": rep64["backward_error"],
           "speedup_mixed_vs_fp64": round(dt64 / dt, 3)}
    log(f"fp64: {dt64 * 1e3:.1f} ms, {out['fp64_lu_tflops']} TFLOP/s, "
        f"bwd err {rep64['backward_error']:.2e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--precision", default="fp16x3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline solve and the plain FP64 LU beside it")
    ap.add_argument("--legs", default="fp64,householder,dense,batched",
                    help="comma list of the reported workloads beside the headline "
                         "(fp64, householder, dense, batched): a profiler run can take one")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device (a kernel of known traffic, to "
                         "calibrate hardware byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the factor-and-solve path runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import torch
    import eigen_amd as ea
    from eigen_amd import mixed

    torch.cuda.set_device(0)
    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": src.numel() * 4, "launches": 3}))
        return
    n = args.n
    g = torch.Generator(device="cuda")
    g.manual_seed(16384)
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g).t()  # col-major
    b = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    torch.cuda.synchronize()
    log(f"inputs ready: n={n}")

    reps = []
    for _ in range(args.warmup):
        x, rep = mixed.gesv_ir(A, b, precision=args.precision)
        log(f"warmup: factor {rep['factor_ms']:.1f} ms refine {rep['refine_ms']:.1f} ms "
            f"bwd err {rep['backward_error']:.2e} ir {rep['ir_iterations']} "
            f"gmres {rep['gmres_iterations']}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x, rep = mixed.gesv_ir(A, b, precision=args.precision)
        reps.append(rep)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    # one more step, untimed, with per-phase HIP events (the events add
    # markers between kernels, so they stay out of the timed steps)
    h = ea.handle()
    ea.lib().eamd_profile_enable(h, 1)
    x, rep = mixed.gesv_ir(A, b, precision=args.precision)
    reps.append(rep)
    torch.cuda.synchronize()
    upd_ms, upd_bytes, upd_flops, upd_cnt = profile_phase(ea, 3)
    phases = {}
    for name, ph in (("panel", 0), ("laswp", 1), ("trsm", 2), ("update", 3), ("pack", 4),
                     ("trsv", 5), ("residual", 6), ("next_panel", 10)):
        ms, by, fl, cnt = profile_phase(ea, ph)
        phases[name] = round(ms, 3)
    ea.lib().eamd_profile_enable(h, 0)
    worst_be = max(r["backward_error"] for r in reps)
    tflops = hpl_flops(n) / dt / 1e12
    log(f"mixed {args.precision}: {dt * 1e3:.1f} ms/step, {tflops:.2f} TFLOP/s, "
        f"bwd err {worst_be:.2e}, phases(ms/step) {phases}")

    extras = {}
    legs = set(args.legs.split(",")) if args.legs else set()
    if args.no_extras:
        legs &= {"fp64"}
    if "fp64" in legs:
        extras.update(fp64_fields(A, b, n, dt))

    if "householder" in legs:
        try:
            Ah = householder_matrix_gpu(n, 64, 5)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            xh, reph = mixed.gesv_ir(Ah, b, precision=args.precision, check_converged=False)
            torch.cuda.synchronize()
            dth = time.perf_counter() - t2
            extras["householder_kappa1e6"] = {
                "tflops": round(hpl_flops(n) / dth / 1e12, 3), "ms": round(dth * 1e3, 1),
                "backward_error": reph["backward_error"], "converged": reph["converged"],
                "ir_iterations": reph["ir_iterations"],
                "gmres_iterations": reph["gmres_iterations"]}
            log(f"householder: {extras['householder_kappa1e6']}")
            del Ah, xh
        except Exception as e:  # reported, not fatal for the headline
            extras["householder_kappa1e6"] = {"error": str(e)[:200]}
    if "dense" in legs:
        try:
            extras.update(dense_fields())
        except Exception as e:
            extras["dense"] = {"error": str(e)[:200]}
    if "batched" in legs:
        extras.update(batched_fields())

    ach_gbs = upd_bytes / (upd_ms * 1e-3) / 1e9 if upd_ms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(ach_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(ach_gbs / PEAK_HBM_GBS, 4),
            "traffic": measured_traffic() if args.n == 16384 and
            args.precision == "fp16x3" else None,
            "kernel": f"mixed_update_kernel ({args.precision})",
            "launches": int(upd_cnt),
            "avg_launch_us": round(upd_ms * 1e3 / max(1, upd_cnt), 2),
            "mfma_tflops": round(upd_flops / (upd_ms * 1e-3) / 1e12, 1) if upd_ms > 0 else 0.0}
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline()
        except Exception as e:
            cpu = {"error": str(e)[:200]}
    out = {
        "metric": METRIC,
        "value": round(tflops, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16x3+f32+f64",
        "data": "synthetic",
        "config": {"workload": f"mixed-precision LU-IR, n={n}, A iid N(0,1), 1 RHS, "
                               f"{args.precision} trailing update, FP32 panel, FP64 "
                               f"refinement to 1e-14",
                   "n": n, "precision": args.precision, "rhs": 1},
        "backward_error": worst_be,
        "ir_iterations": reps[-1]["ir_iterations"],
        "gmres_iterations": reps[-1]["gmres_iterations"],
        "phase_ms_per_step": phases,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
