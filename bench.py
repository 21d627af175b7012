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
    return {"value": round(hpl_flops(n) / dt / 1e12, 5), "unit": "TFLOP/s", "cores": cores,
            "kind": kind,
            "sample": f"PartialPivLU<MatrixXd> factor+solve, n={n} FP64 N(0,1), 1 RHS, "
                      f"{dt:.2f} s (HPL flop count)"}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--precision", default="fp16x3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
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
    # plain FP64 GPU LU (factor + solve) on the same input
    x64, rep64 = mixed.gesv_ir(A, b, precision="fp64")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    x64, rep64 = mixed.gesv_ir(A, b, precision="fp64")
    torch.cuda.synchronize()
    dt64 = time.perf_counter() - t1
    extras["fp64_lu_tflops"] = round(hpl_flops(n) / dt64 / 1e12, 3)
    extras["fp64_lu_ms"] = round(dt64 * 1e3, 2)
    extras["fp64_backward_error"] = rep64["backward_error"]
    extras["speedup_mixed_vs_fp64"] = round(dt64 / dt, 3)
    log(f"fp64: {dt64 * 1e3:.1f} ms, {extras['fp64_lu_tflops']} TFLOP/s, "
        f"bwd err {rep64['backward_error']:.2e}")
    del x64

    if not args.no_extras:
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
        try:
            extras.update(dense_fields())
        except Exception as e:
            extras["dense"] = {"error": str(e)[:200]}
        try:
            from eigen_amd import batched
            extras.update(batched.bench_fields())
        except Exception as e:
            extras["batched"] = {"error": str(e)[:200]}

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
