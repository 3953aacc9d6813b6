#!/usr/bin/env python3
"""Benchmark of the batched dense linear-algebra path on one MI355X.

Default workload (the headline metric): mixed-precision batched Cholesky
solve, k = 8192 SPD 128x128 matrices A = Q diag(lambda) Q^T with lambda
log-spaced on [1, 1e4], 4 right-hand sides each, to backward error 1e-12.
A "step" is one call of gputils_dposv_mixed_batched over the whole batch
(inputs resident in HBM).  The plain all-FP64 kernels are timed on the same
systems (API-shaped factorise + solve, and the fused factor+solve) and
reported beside it, with the speedup of the mixed solver over each.

Prints ONE JSON line (rank 0).  `roofline` is the dominant kernel's
algorithmic bytes over its launch time, the mean of at least 50 HIP-event
timings on the launch stream (the statistic rocprofv3's kernel-trace
summary reports as its average; the median is given beside it).
`cpu_baseline` times a LAPACK batch loop (oracle/lapack_loop.c over the
OpenBLAS that scipy bundles, one thread per call, OpenMP over the matrices)
on the host cores; where that library is missing, the FP64 oracle
(oracle/oracle.c, an unblocked LAPACK restatement) instead, labelled so.

    python bench.py [--steps K] [--warmup W] [--workload NAME]
The default workload, "all", prints the headline line with every other
configuration's measurement under "configs" (gemm: 16384 64x64 FP32
products in both modes; cholesky_small: FP64 32x32 factor + solve;
ls_ragged: 4096 uneven least-squares systems; svd_nullspace: 2048 48x96
nullspace builds and graph-replayed projections), each with its own
roofline and CPU baseline.  --workload NAME prints one configuration's
line alone (copy_calibration: a 1 GiB copy for scaling rocprof byte
counters).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, AMD public spec
FP32_MFMA_PEAK_TFS = 157.3
BF16_MFMA_PEAK_TFS = 2500.0
FP64_PEAK_TFS = 78.6
MIN_EVENT_REPS = 50


def _args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all",
                    choices=["all", "cholesky_mixed", "gemm", "cholesky_small", "ls_ragged", "svd_nullspace",
                             "copy_calibration"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ragged-shapes", default=None,
                    help="ls_ragged: launch shapes per bucket as METHOD:WAVES:PER_CU:QUEUE,... (largest "
                         "bucket first) applied instead of autotuning; the default-shape timing is then skipped")
    ap.add_argument("--ragged-concurrent", type=int, choices=[0, 1], default=None,
                    help="ls_ragged with --ragged-shapes: run the buckets concurrently (1) or in turn (0)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparison runs")
    return ap.parse_args(argv)


def _cpu_threads() -> int:
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def _events_time_ms(fn, reps, setup=None):
    """Per-call time of fn() with HIP events on the current stream; setup()
    runs outside the timed windows.  Returns the list of ms values."""
    import torch
    out = []
    for _ in range(max(reps, MIN_EVENT_REPS)):
        if setup is not None:
            setup()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        e.synchronize()
        out.append(s.elapsed_time(e))
    return out


def _stats(ms):
    return {"mean": statistics.fmean(ms), "median": statistics.median(ms), "min": min(ms), "reps": len(ms)}


def backward_error(A, X, B):
    R = B - np.einsum("kij,krj->kri", A, X, optimize=True)
    an = np.abs(A).sum(axis=2).max(axis=1)[:, None]
    return float(np.max(np.abs(R).max(axis=2) / (an * np.abs(X).max(axis=2) + np.abs(B).max(axis=2))))


# ------------------------------------------------------------ workloads --
def bench_cholesky_mixed(a):
    import torch
    from gputils_amd import api, workloads

    k, n, nrhs, tol = 8192, 128, 4, 1e-12
    t0 = time.time()
    A, B = workloads.cholesky_mixed(k=k, n=n, nrhs=nrhs)
    gen_s = time.time() - t0
    print(f"[bench] generated workload in {gen_s:.1f}s", file=sys.stderr, flush=True)
    dA = torch.from_numpy(A).cuda()
    dB = torch.from_numpy(B).cuda()
    X = torch.empty_like(dB)
    iters = torch.empty(k, dtype=torch.int32, device="cuda")
    info = torch.empty(k, dtype=torch.int32, device="cuda")

    def step():
        api.posv_mixed_batched(dA, dB, X, tol=tol, max_iter=30, iters=iters, info=info)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / a.steps
    value = k * nrhs * a.steps / wall

    # live per-launch timing of the dominant kernel (events on the launch stream)
    ev = _events_time_ms(step, MIN_EVENT_REPS)
    ev_ms = statistics.fmean(ev)
    alg_bytes = k * (n * (n + 1) // 2 * 8 + 2 * n * nrhs * 8)
    alg_flops = k * (n ** 3 / 3 + n ** 2 / 2 + n / 6 + 2 * n * n * nrhs)

    Xh = X.cpu().numpy()
    it = iters.cpu().numpy()
    inf = info.cpu().numpy()
    berr = backward_error(A, Xh, B)
    print(f"[bench] mixed: {ms_per_step:.3f} ms/step, events {ev_ms:.3f} ms, berr {berr:.2e}, "
          f"iters {np.bincount(np.maximum(it, 0)).tolist()}", file=sys.stderr, flush=True)

    plain = {}
    if not a.no_plain:
        # API-shaped plain FP64: potrf (writes L) then potrs; A refreshed outside the window
        dL = torch.empty_like(dA)
        dX = torch.empty_like(dB)

        def refresh():
            dL.copy_(dA)
            dX.copy_(dB)

        def api_shaped():
            api.potrf_batched(dL, info)
            api.potrs_batched(dL, dX, info)

        def fused():
            api.posv_batched(dL, dX, info)

        # algorithmic bytes: fused reads the lower triangle of A and B and
        # writes L and X; API-shaped also reads L back for the solve
        tri = n * (n + 1) // 2 * 8
        plain_bytes = {"fused": k * (2 * tri + 2 * n * nrhs * 8),
                       "api_shaped": k * (3 * tri + 3 * n * nrhs * 8)}
        plain_kernel = {"fused": "posv_kernel<double,8,true>", "api_shaped": "potrf_kernel<double,8>+potrs_kernel<double,8>"}
        for name, fn in (("api_shaped", api_shaped), ("fused", fused)):
            _events_time_ms(fn, 2, refresh)
            tl = _events_time_ms(fn, MIN_EVENT_REPS, refresh)
            t = statistics.fmean(tl)
            refresh()
            fn()
            torch.cuda.synchronize()
            plain[name] = {
                "ms_per_step": round(t, 4),
                "event_ms": _stats(tl),
                "solves_per_s": k * nrhs / (t * 1e-3),
                "backward_error": backward_error(A, dX.cpu().numpy(), B),
                "speedup_of_mixed": round(t / ev_ms, 3),
                "roofline": _roofline("hbm", plain_bytes[name], t, plain_kernel[name],
                                      "lower triangle of A read (and of L written, and read again when "
                                      "API-shaped), B read and X written, 8 B each"),
                "fp64_rate_tflops": alg_flops / (t * 1e-3) / 1e12,
                "fp64_frac": alg_flops / (t * 1e-3) / 1e12 / FP64_PEAK_TFS,
            }
            print(f"[bench] plain {name}: {t:.3f} ms, berr {plain[name]['backward_error']:.2e}",
                  file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline_cholesky(A, B)

    achieved = alg_bytes / (ev_ms * 1e-3) / 1e9
    ev_stats = _stats(ev)
    out = {
        "metric": "batched solves/s, mixed vs plain FP64 at equal 1e-12 residual; % HBM/MFMA peak",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded SplitMix64)",
        "config": {"workload": "cholesky_mixed: k=8192 SPD 128x128, lambda=10^(4j/127), 4 RHS, "
                               "FP32 factor + FP64 refinement to backward error 1e-12",
                   "k": k, "n": n, "nrhs": nrhs, "tol": tol},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": _traffic("posv_mixed_wave_kernel<8>"),
                     "kernel": "posv_mixed_wave_kernel<8> (+ FP64 fallback scan launch)",
                     "algorithmic_bytes": alg_bytes, "algorithmic_note": "lower triangle of A read once "
                     "(n(n+1)/2 * 8 B), B read and X written (n * nrhs * 8 B each), per system",
                     "event_ms": ev_ms, "event_ms_stats": ev_stats,
                     "fp_rate_tflops": alg_flops / (ev_ms * 1e-3) / 1e12},
        "cpu_baseline": cpu,
        "accuracy": {"max_backward_error": berr, "fp64_fallbacks": int(np.sum(it < 0)),
                     "refinement_steps_histogram": np.bincount(np.maximum(it, 0)).tolist(),
                     "info_nonzero": int(np.sum(inf != 0))},
        "plain_fp64": plain,
    }
    if plain:
        # like for like: HIP-event medians of one call on both sides (the
        # mixed call includes its FP64 fallback scan launch)
        best_ms = min(p["ms_per_step"] for p in plain.values())
        out["speedup_vs_best_plain_fp64"] = best_ms / ev_ms
        out["speedup_vs_best_plain_fp64_wall"] = value / max(p["solves_per_s"] for p in plain.values())
    return out


def _lapack():
    """The LAPACK batch loops (oracle/lapack_loop.py), or None."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lapack_loop
    return lapack_loop if lapack_loop.lib() is not None else None


def _lapack_kind(ll):
    return {"kind": "lapack", "library": f"OpenBLAS LAPACK bundled with scipy ({ll.library_name()}), one "
                                         "thread per call, OpenMP over the matrices"}


def _repeat(fn, min_seconds):
    """Runs fn() until min_seconds of timed wall clock; returns (passes, seconds)."""
    fn()  # warm-up
    passes, dt = 0, 0.0
    while dt < min_seconds:
        t1 = time.perf_counter()
        fn()
        dt += time.perf_counter() - t1
        passes += 1
    return passes, dt


def cpu_baseline_cholesky(A, B, min_seconds=2.0):
    """LAPACK dpotrf + dpotrs over the same systems (whole passes over the
    batch, repeated until min_seconds of wall time); the FP64 oracle
    (dpotrf + dpotrs restated, unblocked) where LAPACK is missing."""
    threads = _cpu_threads()
    ll = _lapack()
    if ll is not None:
        bad = []
        passes, dt = _repeat(lambda: bad.append(ll.posv(A, B, threads)[1]), min_seconds)
        return {"value": passes * A.shape[0] * B.shape[1] / dt, "unit": "solves/s", "cores": threads,
                **_lapack_kind(ll),
                "sample": f"{passes} pass(es) over all {A.shape[0]} systems (dpotrf + dpotrs, {B.shape[1]} RHS), "
                          f"{dt:.2f}s of timed wall clock, {bad[-1]} failures"}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    pyoracle.posv_batched(A[:64].copy(), B[:64].copy(), threads)   # warm-up
    solved, bad, passes, dt = 0, 0, 0, 0.0
    while dt < min_seconds:
        As, Bs = A.copy(), B.copy()          # the factorisation works in place
        t1 = time.perf_counter()
        bad += pyoracle.posv_batched(As, Bs, threads)
        dt += time.perf_counter() - t1
        solved += A.shape[0] * B.shape[1]
        passes += 1
    return {"value": solved / dt, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{passes} pass(es) over all {A.shape[0]} systems (FP64 oracle factor + "
                      f"{B.shape[1]}-RHS solve), {dt:.2f}s of timed wall clock, {bad} failures"}


def _line(metric, value, unit, a, ms_per_step, dtype, config, roofline, cpu, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded SplitMix64)", "config": config,
           "roofline": roofline, "cpu_baseline": cpu}
    out.update(extra)
    return out


def _roofline(bound, alg, ev_ms, kernel, unit_bytes_note, traffic_key=None):
    """bound "hbm" (alg in bytes) or (name, peak TFLOP/s) (alg in flops)."""
    if bound == "hbm":
        achieved = alg / (ev_ms * 1e-3) / 1e9
        peak, unit = HBM_PEAK_GBS, "GB/s"
    else:
        achieved = alg / (ev_ms * 1e-3) / 1e12
        peak, unit = bound[1], "TFLOP/s"
        bound = bound[0]
    return {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
            "traffic": _traffic(traffic_key or kernel), "kernel": kernel, "algorithmic": alg, "algorithmic_note": unit_bytes_note,
            "event_ms": ev_ms}


def _traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    or None when no summary exists for it.  A name ending in '*' sums every
    kernel whose name starts with the rest (one launch of each per step)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    kernels = tab.get("kernels", {})
    if "+" in kernel:  # one launch of each listed kernel per step
        vals = [_traffic(k) for k in kernel.split("+")]
        return sum(vals) if None not in vals else None
    if kernel.endswith("*"):
        vals = [v.get("hbm_bytes_per_launch") for k, v in kernels.items() if k.startswith(kernel[:-1])]
        return sum(vals) if vals and None not in vals else None
    rec = kernels.get(kernel)
    return None if rec is None else rec.get("hbm_bytes_per_launch")


def _timed_steps(a, step):
    import torch
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def bench_gemm(a):
    """Batched C = A B, 16384 64x64 FP32 matrices: FP32 MFMA kernel (the
    step) and the split-BF16 3-term mode, errors vs FP64 host GEMM."""
    import torch
    from gputils_amd import api, workloads
    A, B = workloads.gemm()
    k, n = A.shape[0], A.shape[1]
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dC = torch.empty_like(dA)
    # FP64 host product of every matrix (tensor layout: C^T = B^T A^T)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    ref = np.matmul(B64, A64)
    absAB = np.matmul(np.abs(B64), np.abs(A64))
    del A64, B64
    alg = 3 * k * n * n * 4
    modes = {}
    kernels = {"fp32": "gemm_staged_kernel<0>", "bf16x3": "gemm_staged_kernel<1>"}
    for mode in ("fp32", "bf16x3"):
        fn = lambda: api.gemm_batched(dA, dB, dC, mode=mode)
        wall = _timed_steps(a, fn)
        evl = _events_time_ms(fn, MIN_EVENT_REPS)
        ev = statistics.fmean(evl)
        err = np.abs(dC.cpu().numpy().astype(np.float64) - ref)
        modes[mode] = {"ms_per_step": 1e3 * wall / a.steps, "event_ms": ev, "event_ms_stats": _stats(evl),
                       "GB/s": alg / (ev * 1e-3) / 1e9, "TFLOP/s": 2 * k * n ** 3 / (ev * 1e-3) / 1e12,
                       "roofline": _roofline("hbm", alg, ev, kernels[mode], "A+B read, C written: 3*64*64*4 B "
                                             "per matrix"),
                       "max_err_rel_absAB": float((err / absAB).max()),
                       "max_err_rel_C": float(err.max() / np.abs(ref).max()),
                       "error_over": f"all {k} products vs FP64 host GEMM"}
        print(f"[bench] gemm {mode}: {ev:.4f} ms {modes[mode]['GB/s']:.0f} GB/s, max |dC|/(|A||B|) "
              f"{modes[mode]['max_err_rel_absAB']:.2e}", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu_baseline:
        cpu = _cpu_gemm(A, B)
    m = modes["fp32"]
    return _line("batched GEMMs/s (C = A B, 64x64x64, FP32)", k / (m["ms_per_step"] * 1e-3), "GEMMs/s", a,
                 m["ms_per_step"], "f32",
                 {"workload": "gemm: k=16384 64x64 by 64x64 FP32, U[-1,1]", "k": k, "m": n, "n": n, "kk": n},
                 m["roofline"], cpu, modes=modes)


def _cpu_gemm(A, B, min_seconds=2.0):
    threads = _cpu_threads()
    ll = _lapack()
    if ll is not None:
        passes, dt = _repeat(lambda: ll.sgemm(A, B, threads), min_seconds)
        return {"value": passes * A.shape[0] / dt, "unit": "GEMMs/s", "cores": threads, **_lapack_kind(ll),
                "sample": f"{passes} pass(es) of sgemm over all {A.shape[0]} products, {dt:.2f}s timed"}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    sample = 2048
    done, dt = 0, 0.0
    while dt < min_seconds:
        t1 = time.perf_counter()
        pyoracle.gemm_batched(A[:sample], B[:sample], threads)
        dt += time.perf_counter() - t1
        done += sample
    return {"value": done / dt, "unit": "GEMMs/s", "cores": threads, "kind": "port",
            "sample": f"{done} FP32 64x64x64 products (first {sample} matrices, repeated) by the FP64 oracle, "
                      f"{dt:.2f}s timed"}


def bench_cholesky_small(a):
    """FP64 factor + solve of 32x32 SPD systems: k = 1024 (the parity case)
    and k = 131072 (fills the GPU; used for the roofline)."""
    import torch
    from gputils_amd import api, workloads
    res = {}
    for k in (1024, 131072):
        A, b = workloads.cholesky_small(k=k)
        dA0, db0 = torch.from_numpy(A).cuda(), torch.from_numpy(b).cuda()
        dA, db = dA0.clone(), db0.clone()
        info = torch.empty(k, dtype=torch.int32, device="cuda")

        def refresh():
            dA.copy_(dA0)
            db.copy_(db0)

        fn = lambda: api.posv_batched(dA, db, info)
        _events_time_ms(fn, a.warmup, refresh)
        tl = _events_time_ms(fn, MIN_EVENT_REPS, refresh)
        t = statistics.fmean(tl)
        refresh()
        fn()
        x = db.cpu().numpy()
        resid = float(np.max(np.linalg.norm(b[:, 0, :] - np.einsum("kij,kj->ki", A, x[:, 0, :]), axis=1) /
                             np.linalg.norm(b[:, 0, :], axis=1)))
        alg = k * (2 * (32 * 33 // 2) * 8 + 2 * 32 * 8)
        res[str(k)] = {"event_ms": t, "event_ms_stats": _stats(tl), "solves_per_s": k / (t * 1e-3),
                       "GB/s": alg / (t * 1e-3) / 1e9, "max_rel_residual": resid}
        print(f"[bench] cholesky_small k={k}: {t:.4f} ms", file=sys.stderr, flush=True)
    big = res["131072"]
    cpu = None
    if not a.no_cpu_baseline:
        A, b = workloads.cholesky_small(k=1024)
        cpu = cpu_baseline_cholesky(A, b)
    return _line("batched FP64 solves/s (32x32 factor + solve)", big["solves_per_s"], "solves/s", a,
                 big["event_ms"], "f64",
                 {"workload": "cholesky_small: k=131072 (and 1024) SPD 32x32, A = M M^T + 32 I, 1 RHS",
                  "k": 131072, "n": 32},
                 _roofline("hbm", 131072 * (2 * (32 * 33 // 2) * 8 + 2 * 32 * 8), big["event_ms"],
                           "posv_small_kernel<double,32,true,false>",
                           "lower triangle of A read and of L written, b read and x written (8 B each)"),
                 cpu, results=res)


def bench_ls_ragged(a):
    """4096 uneven least-squares systems in one plan execution."""
    import torch
    from gputils_amd import api, workloads
    ms, ns, offA, offB, Af, Bf = workloads.ls_ragged()
    plan = api.RaggedLeastSquares(ms, ns, offA, offB)
    dA = torch.from_numpy(Af).cuda()
    dB0 = torch.from_numpy(Bf).cuda()
    dB = dB0.clone()

    def refresh():
        dB.copy_(dB0)

    count = len(ms)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    rng = np.random.default_rng(0)
    sample64 = rng.choice(count, 64, replace=False)
    refs = {}
    for s in sample64:
        m, n = int(ms[s]), int(ns[s])
        Am = Af[offA[s]:offA[s] + m * n].reshape(n, m).T
        refs[s] = pyoracle.gels(Am, Bf[offB[s]:offB[s] + m]).ravel()

    def run(steps):
        fn = lambda: plan(dA, dB, refine_steps=steps)
        _events_time_ms(fn, a.warmup, refresh)
        t = statistics.fmean(_events_time_ms(fn, MIN_EVENT_REPS, refresh))
        refresh()
        fn()
        out = dB.cpu().numpy()
        errs = [np.linalg.norm(out[offB[s]:offB[s] + int(ns[s])] - refs[s]) / np.linalg.norm(refs[s])
                for s in sample64]
        print(f"[bench] ls_ragged, {steps} refinement step(s): {t:.4f} ms, max rel err {max(errs):.2e}",
              file=sys.stderr, flush=True)
        return t, errs

    # the configuration's method: FP32 QR + one FP64 refinement step; two
    # steps (the API default) reported beside it.  The plan's launch shapes
    # are autotuned once on these inputs (plan-time work, like building the
    # plan, outside the timed region); the default shapes are timed first.
    default_shapes = plan.describe()
    if a.ragged_shapes:
        given = [tuple(int(v) for v in x.split(":")) for x in a.ragged_shapes.split(",")]
        if len(given) != len(default_shapes):
            raise SystemExit(f"--ragged-shapes needs {len(default_shapes)} shapes, got {len(given)}")
        for b, (method, waves, per_cu, queue) in enumerate(given):
            plan.configure(b, method, waves, per_cu, queue)
        if a.ragged_concurrent is not None:
            plan.concurrent = bool(a.ragged_concurrent)
        ev_default, t_tune, shapes, default_concurrent = None, None, plan.describe(), None
        print(f"[bench] ls_ragged given shapes {shapes}", file=sys.stderr, flush=True)
    else:
        ev_default, _ = run(1)
        default_concurrent = plan.concurrent
        torch.cuda.synchronize()
        t_tune = time.perf_counter()
        shapes = plan.tune(dA, dB0, refine_steps=1)
        t_tune = time.perf_counter() - t_tune
        print(f"[bench] ls_ragged autotune: {t_tune:.2f}s, shapes {shapes}", file=sys.stderr, flush=True)
    ev, errs = run(1)
    ev2, errs2 = run(2)
    alg = int(8 * (np.sum(ms.astype(np.int64) * ns) + np.sum(ms) + np.sum(ns)))
    flops = float(np.sum(2.0 * ms * ns.astype(np.float64) ** 2 - 2.0 * ns.astype(np.float64) ** 3 / 3 +
                         4.0 * ms * ns - ns.astype(np.float64) ** 2))
    cpu = None
    if not a.no_cpu_baseline:
        threads = _cpu_threads()
        ll = _lapack()
        if ll is not None:
            passes, dt = _repeat(lambda: ll.gels_ragged(ms, ns, offA, offB, Af, Bf, threads), 2.0)
            cpu = {"value": passes * count / dt, "unit": "systems/s", "cores": threads, **_lapack_kind(ll),
                   "sample": f"{passes} pass(es) of dgels over all {count} systems, {dt:.2f}s timed"}
        else:
            sample = 1024
            Bs = Bf.copy()
            t1 = time.perf_counter()
            pyoracle.gels_ragged(ms[:sample], ns[:sample], offA[:sample], offB[:sample], Af.copy(), Bs, threads)
            dt = time.perf_counter() - t1
            cpu = {"value": sample / dt, "unit": "systems/s", "cores": threads, "kind": "port",
                   "sample": f"first {sample} of the 4096 systems (FP64 Householder dgels restated), {dt:.2f}s"}
    return _line("ragged least-squares systems/s (4096 systems, m in [16,512], n in [8,64])",
                 count / (ev * 1e-3), "systems/s", a, ev, "f32+f64",
                 {"workload": "ls_ragged: 4096 systems, n~U{8..64}, m~U{max(16,n)..512}, N(0,1), FP32 QR + "
                              "one FP64 augmented-system refinement step", "count": count, "refine_steps": 1},
                 _roofline("hbm", alg, ev, "gels_ragged_kernel<*> (one launch per LDS-size bucket)",
                           "A read once, b read, x written (8 B each)", traffic_key="gels_ragged_kernel<*"),
                 cpu, accuracy={"max_rel_err_vs_dgels_sample64": float(max(errs)),
                                "median_rel_err_vs_dgels_sample64": float(np.median(errs))},
                 two_refinement_steps={"event_ms": ev2, "systems_per_s": count / (ev2 * 1e-3),
                                       "max_rel_err_vs_dgels_sample64": float(max(errs2))},
                 autotune={"tuned": not a.ragged_shapes, "tune_seconds": t_tune,
                           "default_shape_event_ms": ev_default,
                           "default_shapes": [[d["maxe"], d["method"], d["waves"], d["per_cu"], d["grid"], d["queue"]]
                                              for d in default_shapes],
                           "shapes": [[d["maxe"], d["method"], d["waves"], d["per_cu"], d["grid"], d["queue"]]
                                      for d in shapes],
                           "shape_fields": ["maxe", "method", "waves", "per_cu", "grid", "queue"],
                           "default_concurrent_buckets": default_concurrent, "concurrent_buckets": plan.concurrent},
                 fp_rate_tflops=flops / (ev * 1e-3) / 1e12)


def bench_svd_nullspace(a):
    """2048 rank-40 48x96 FP32 matrices: (i) SVD + nullspace build, (ii)
    1000 project calls replayed from one HIP graph (the step)."""
    import torch
    from gputils_amd import api, workloads
    A_t, x = workloads.svd_nullspace()
    k, n, m = A_t.shape
    dA = torch.from_numpy(A_t).cuda()
    build = lambda: api.nullspace_batched(dA)
    for _ in range(2):
        build()
    torch.cuda.synchronize()
    tb = statistics.fmean(_events_time_ms(build, MIN_EVENT_REPS))
    N, P, rank = build()
    At = api.transpose_batched(dA)
    # the SVD as the nullspace build runs it (no right singular vectors:
    # sweeps on R^T), and the full SVD with VT for reference
    S, U, VT, info, sw = api.gesvd_batched(At, compute_u=True, sweeps=True, compute_v=False)
    tsvd_l = _events_time_ms(lambda: api.gesvd_batched(At, compute_u=True, compute_v=False), MIN_EVENT_REPS)
    tsvd = statistics.fmean(tsvd_l)
    _, _, _, _, sw_full = api.gesvd_batched(At, compute_u=True, sweeps=True)
    tsvd_full = statistics.fmean(_events_time_ms(lambda: api.gesvd_batched(At, compute_u=True), MIN_EVENT_REPS))
    torch.cuda.synchronize()
    # the same SVD with every sweep in FP64, for comparison
    prev = api.set_fp32_sweeps(False)
    try:
        _, _, _, _, sw_64 = api.gesvd_batched(At, compute_u=True, sweeps=True, compute_v=False)
        tsvd_64 = statistics.fmean(_events_time_ms(lambda: api.gesvd_batched(At, compute_u=True, compute_v=False),
                                                   MIN_EVENT_REPS))
    finally:
        api.set_fp32_sweeps(prev)
    torch.cuda.synchronize()
    # flops the SVD kernel executes per matrix (A^T is mm x nn, mm = 96,
    # nn = 48): Householder QR 2 mm nn^2 - 2 nn^3 / 3 (FP64); each Jacobi
    # sweep on the nn x nn triangle, every pair counted as rotated: three dot
    # products and the rotation of two W and two V columns, 18 nn flops per
    # pair, nn (nn-1)/2 pairs (FP32 or FP64 as the sweep runs); the step
    # between them, F = X X^T, X - F X / 2 and W0 X (lower triangular W0),
    # 5 nn^3 flops in FP64; U = Q diag(U_R, I): 4 (mm - j) mm flops per
    # reflector j (FP64).
    mm, nn = n, m
    s64 = (sw & 0xFFFF).cpu().numpy().astype(np.float64)
    s32 = (sw >> 16).cpu().numpy().astype(np.float64)
    qr_f = 2 * mm * nn ** 2 - 2 * nn ** 3 / 3
    sweep_f = 18 * nn * nn * (nn - 1) / 2
    step_f = 5 * nn ** 3
    u_f = sum(4 * (mm - j) * mm for j in range(nn))
    f64 = float(np.sum(qr_f + s64 * sweep_f + (s32 > 0) * step_f + u_f))
    f32 = float(np.sum(s32 * sweep_f))
    # the time these flops need at the FP64 and FP32 vector peaks (FP32
    # without packed instructions runs at the FP64 rate, 78.6 TFLOP/s)
    svd_roof = _roofline(("fp64 + fp32 vector", FP64_PEAK_TFS), f64 + f32, tsvd, "gesvd_kernel<float,3,true>",
                         "flops executed: FP64 Householder QR of the 96x48 A^T, the measured FP32 then FP64 Jacobi "
                         "sweeps on the 48x48 R^T (18*48 flops per pair, all 1128 pairs counted), the FP64 step "
                         "between them (5*48^3), FP64 U = Q diag(U_R, I); the peak is FP64's, which non-packed FP32 "
                         "arithmetic shares")
    svd_roof["fp64_flops"] = f64
    svd_roof["fp32_flops"] = f32
    svd_roof["sweeps_fp32_mean"] = float(s32.mean())
    svd_roof["sweeps_fp64_max"] = int(s64.max())
    svd_roof["sweeps_fp64_mean"] = float(s64.mean())
    svd_roof["all_fp64_sweeps"] = {"gesvd_ms": tsvd_64, "sweeps_max": int((sw_64 & 0xFFFF).max()),
                                   "sweeps_mean": float((sw_64 & 0xFFFF).double().mean()),
                                   "speedup_of_fp32_first": tsvd_64 / tsvd}
    reps = 1000

    def capture(call):
        """One HIP graph of `reps` project calls; call(i) makes call i."""
        st = torch.cuda.Stream()
        st.wait_stream(torch.cuda.current_stream())
        gr = torch.cuda.CUDAGraph()
        with torch.cuda.stream(st):
            call(0)
            torch.cuda.synchronize()
            with torch.cuda.graph(gr, stream=st):
                for i in range(reps):
                    call(i)
        torch.cuda.synchronize()
        return gr

    # Nullspace::project as N (N^T b) from the nullity columns of N (the
    # library's form) and, for comparison, as the product P b
    b = torch.from_numpy(x).cuda()
    replays = max(5, min(a.steps, 50))
    a_graph = argparse.Namespace(**{**vars(a), "steps": replays})

    def graph_rate(call):
        gr = capture(call)
        b.copy_(torch.from_numpy(x))
        w = _timed_steps(a_graph, gr.replay)
        del gr
        return w, 1e3 * w / (replays * reps)

    wall, per_call_ms = graph_rate(lambda i: api.nullspace_project(N, rank, b))
    ev = statistics.fmean(_events_time_ms(lambda: api.nullspace_project(N, rank, b), MIN_EVENT_REPS))
    # the same calls with N cycled through copies whose read parts exceed
    # 256 MB in all, so that no call finds its N in the Infinity Cache
    nul = n - rank.to(torch.int64).clamp(max=n)
    n_bytes = int(nul.sum()) * n * N.element_size()
    copies = int(np.ceil(300e6 / n_bytes))
    Ns = [N.clone() for _ in range(copies)]
    _, per_call_cycled_ms = graph_rate(lambda i: api.nullspace_project(Ns[i % copies], rank, b))
    del Ns
    _, per_call_p_ms = graph_rate(lambda i: api.project(P, b))
    copies_p = int(np.ceil(300e6 / (P.numel() * P.element_size())))
    Ps = [P.clone() for _ in range(copies_p)]
    _, per_call_p_cycled_ms = graph_rate(lambda i: api.project(Ps[i % copies_p], b))
    del Ps
    # accuracy on the projected vectors of a fresh call
    b.copy_(torch.from_numpy(x))
    api.nullspace_project(N, rank, b)
    z = b.cpu().double().numpy()[:, 0, :]
    Am = np.swapaxes(A_t.astype(np.float64), 1, 2)
    res = np.linalg.norm(np.einsum("kij,kj->ki", Am, z), axis=1)
    rel = float(np.max(res / np.linalg.norm(x[:, 0, :].astype(np.float64), axis=1)))
    alg = n_bytes + k * 2 * n * 4
    alg_p = k * (n * n * 4 + 2 * n * 4)
    print(f"[bench] svd+nullspace build {tb:.3f} ms (svd {tsvd:.3f} ms, sweeps FP32 {s32.mean():.1f} + FP64 "
          f"{s64.mean():.1f}; all-FP64 sweeps {tsvd_64:.3f} ms), "
          f"{svd_roof['achieved']:.2f} TFLOP/s); project {per_call_ms * 1e3:.2f} us/call in graph "
          f"({per_call_cycled_ms * 1e3:.2f} us with N cycled over {copies} copies), {ev * 1e3:.2f} us eager; "
          f"as P b {per_call_p_ms * 1e3:.2f} / {per_call_p_cycled_ms * 1e3:.2f} us; max |A P x|/|x| {rel:.2e}",
          file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu_baseline:
        threads = _cpu_threads()
        ll = _lapack()
        if ll is not None:
            passes, dt = _repeat(lambda: ll.nullspace(A_t, 1e-6, threads), 2.0)
            Pc = np.ascontiguousarray(api.to_matrices(P).cpu().numpy())
            xc = np.ascontiguousarray(x.copy())
            calls = 100
            t1 = time.perf_counter()
            ll.project(Pc, xc, calls, threads)
            dtp = time.perf_counter() - t1
            # about one second of calls for the reported rate
            calls = int(min(100000, max(100, calls * 1.0 / max(dtp, 1e-6))))
            t1 = time.perf_counter()
            ll.project(Pc, xc, calls, threads)
            dtp = time.perf_counter() - t1
            cpu = {"value": calls / dtp, "unit": "calls/s", "cores": threads, **_lapack_kind(ll),
                   "sample": f"{calls} rounds of sgemv (P_i x_i, all {k} matrices), {dtp:.2f}s",
                   "nullspace_builds_per_s": passes * k / dt,
                   "nullspace_sample": f"{passes} pass(es) of dgesvd (jobu 'A') of A^T + dgemm N N^T over all {k} "
                                       f"matrices, {dt:.2f}s"}
        else:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle
            pyoracle.build()
            sample = 512
            t1 = time.perf_counter()
            pyoracle.nullspace_batched(A_t[:sample].astype(np.float64), 1e-6, threads)
            dt = time.perf_counter() - t1
            cpu = {"value": sample / dt, "unit": "nullspace builds/s", "cores": threads, "kind": "port",
                   "sample": f"SVD (one-sided Jacobi) + N + P of the first {sample} matrices in FP64, {dt:.2f}s"}
    # n = 96 FP32 takes the 8-lane column-group instantiation (3 loads per lane)
    roof = _roofline("hbm", alg, per_call_cycled_ms, "nullspace_project_kernel<float,0,3>",
                     f"the nullity columns of N read (96 x (96 - rank) x 4 B, {n_bytes / 1e6:.1f} MB in all) + b read "
                     f"and written per matrix; calls cycle through {copies} copies of N "
                     f"({copies * n_bytes / 1e6:.0f} MB read > the 256 MB Infinity Cache)")
    roof["via_P"] = {"kernel": "gemv_quad_kernel", "us_per_call": per_call_p_ms * 1e3,
                     "us_per_call_cycled": per_call_p_cycled_ms * 1e3,
                     "achieved_cycled": alg_p / (per_call_p_cycled_ms * 1e-3) / 1e9, "algorithmic": alg_p,
                     "note": "the same projection as the product P b (96 x 96 P per matrix), P cycled over "
                             f"{copies_p} copies"}
    roof["cache_resident"] = {"achieved": alg / (per_call_ms * 1e-3) / 1e9, "unit": "GB/s",
                              "note": "one N for all 1000 calls: served in part from the Infinity Cache, so "
                                      "not an HBM rate"}
    return _line("Nullspace::project calls/s (2048 projections onto 56-dim nullspaces in R^96, 1000 calls per "
                 "HIP graph replay)",
                 1e3 / per_call_ms, "calls/s", a_graph, 1e3 * wall / replays, "f32",
                 {"workload": "svd_nullspace: k=2048 rank-40 48x96 FP32, sigma~U[0.1,10]; nullspace build then "
                              "1000 graph-captured project calls per step", "k": k, "m": m, "n": n,
                  "calls_per_step": reps},
                 roof, cpu,
                 build={"nullspace_build_ms": tb, "gesvd_ms": tsvd,
                        "gesvd_with_vt_ms": tsvd_full, "max_sweeps_with_vt": int((sw_full & 0xFFFF).max()),
                        "ranks": sorted(set(rank.cpu().tolist())), "info_nonzero": int((info != 0).sum()),
                        "gesvd_roofline": svd_roof, "nullspace_builds_per_s": k / (tb * 1e-3)},
                 accuracy={"max_rel_A_Px": rel}, project_eager_event_ms=ev,
                 project_cycled_us_per_call=per_call_cycled_ms * 1e3)


def bench_copy_calibration(a):
    """Plain 1 GiB device copy (torch) for scaling the FETCH_SIZE /
    WRITE_SIZE counters to bytes."""
    import torch
    nbytes = 1 << 30
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    fn = lambda: dst.copy_(src)
    wall = _timed_steps(a, fn)
    ev = statistics.median(_events_time_ms(fn, max(5, a.steps)))
    return _line("device copy GB/s (counter calibration)", 2 * nbytes / (ev * 1e-3) / 1e9, "GB/s", a,
                 1e3 * wall / a.steps, "f32", {"workload": "copy_calibration: 1 GiB read + 1 GiB written"},
                 _roofline("hbm", 2 * nbytes, ev, "copy", "1 GiB read + 1 GiB written"), None)


def main(argv=None):
    a = _args(argv)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        return 2
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        return 1
    from gputils_amd import _lib
    _lib.lib()  # fails loudly when the HIP library is not built
    table = {"cholesky_mixed": bench_cholesky_mixed, "gemm": bench_gemm, "cholesky_small": bench_cholesky_small,
             "ls_ragged": bench_ls_ragged, "svd_nullspace": bench_svd_nullspace,
             "copy_calibration": bench_copy_calibration}
    if a.workload != "all":
        print(json.dumps(table[a.workload](a)), flush=True)
        return 0
    out = bench_cholesky_mixed(a)
    # every other configuration, timed in the same run (one GPU, in turn)
    keep = ("metric", "value", "unit", "ms_per_step", "dtype", "config", "roofline", "cpu_baseline")
    out["configs"] = {}
    for key, name in (("c1_cholesky_small", "cholesky_small"), ("c2_gemm", "gemm"), ("c4_ls_ragged", "ls_ragged"),
                      ("c5_svd_nullspace", "svd_nullspace")):
        print(f"[bench] configuration {name}", file=sys.stderr, flush=True)
        try:
            line = table[name](a)
            out["configs"][key] = {kk: v for kk, v in line.items()
                                   if kk in keep or kk not in ("n_gpus", "steps", "warmup", "higher_is_better",
                                                               "scaling", "vs_baseline", "data")}
        except Exception as e:  # one configuration's failure must not hide the headline
            out["configs"][key] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
