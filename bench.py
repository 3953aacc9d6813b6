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
algorithmic bytes over its launch time measured with HIP events on the
launch stream; `cpu_baseline` times the FP64 oracle (oracle/oracle.c, an
unblocked LAPACK restatement, OpenMP over matrices) on the host cores.

    python bench.py [--steps K] [--warmup W] [--workload NAME]
Other workloads (gemm, cholesky_small) are for profiling and print their own
line; they are not the headline.
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


def _args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cholesky_mixed",
                    choices=["cholesky_mixed", "gemm", "cholesky_small"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
    for _ in range(reps):
        if setup is not None:
            setup()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        e.synchronize()
        out.append(s.elapsed_time(e))
    return out


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
    ev = _events_time_ms(step, max(5, min(a.steps, 20)))
    ev_ms = statistics.median(ev)
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

        for name, fn in (("api_shaped", api_shaped), ("fused", fused)):
            _events_time_ms(fn, 2, refresh)
            t = statistics.median(_events_time_ms(fn, max(5, min(a.steps, 10)), refresh))
            refresh()
            fn()
            torch.cuda.synchronize()
            plain[name] = {
                "ms_per_step": round(t, 4),
                "solves_per_s": k * nrhs / (t * 1e-3),
                "backward_error": backward_error(A, dX.cpu().numpy(), B),
                "speedup_of_mixed": round((k * nrhs * a.steps / wall) / (k * nrhs / (t * 1e-3)), 3),
            }
            print(f"[bench] plain {name}: {t:.3f} ms, berr {plain[name]['backward_error']:.2e}",
                  file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline_cholesky(A, B)

    achieved = alg_bytes / (ev_ms * 1e-3) / 1e9
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "posv_mixed_kernel<8> (+ FP64 fallback scan launch)",
                     "algorithmic_bytes": alg_bytes, "event_ms": ev_ms,
                     "fp_rate_tflops": alg_flops / (ev_ms * 1e-3) / 1e12},
        "cpu_baseline": cpu,
        "accuracy": {"max_backward_error": berr, "fp64_fallbacks": int(np.sum(it < 0)),
                     "refinement_steps_histogram": np.bincount(np.maximum(it, 0)).tolist(),
                     "info_nonzero": int(np.sum(inf != 0))},
        "plain_fp64": plain,
    }
    if plain:
        best = max(p["solves_per_s"] for p in plain.values())
        out["speedup_vs_best_plain_fp64"] = value / best
    return out


def cpu_baseline_cholesky(A, B, min_seconds=2.0):
    """FP64 oracle (dpotrf + dpotrs restated, unblocked) over the same
    systems, OpenMP over matrices; whole passes over the batch are repeated
    until min_seconds of wall time have elapsed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    threads = _cpu_threads()
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
            "sample": f"{passes} pass(es) over all {A.shape[0]} systems (FP64 factor + "
                      f"{B.shape[1]}-RHS solve), {dt:.2f}s of timed wall clock, {bad} failures"}


def bench_gemm(a):
    import torch
    from gputils_amd import api, workloads
    A, B = workloads.gemm()
    k, n = A.shape[0], A.shape[1]
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dC = torch.empty_like(dA)
    res = {}
    for mode in ("fp32", "bf16x3"):
        fn = lambda: api.gemm_batched(dA, dB, dC, mode=mode)
        for _ in range(a.warmup):
            fn()
        t = statistics.median(_events_time_ms(fn, a.steps))
        Cg = dC.cpu().numpy().astype(np.float64)
        ref = np.matmul(np.swapaxes(B[:256].astype(np.float64), 1, 2), np.swapaxes(A[:256].astype(np.float64), 1, 2))
        ref = np.swapaxes(ref, 1, 2)  # C^T = B^T A^T in tensor layout
        err = np.abs(Cg[:256] - np.swapaxes(np.swapaxes(A[:256], 1, 2).astype(np.float64) @
                                            np.swapaxes(B[:256], 1, 2).astype(np.float64), 1, 2))
        absAB = np.swapaxes(np.abs(np.swapaxes(A[:256], 1, 2)).astype(np.float64) @
                            np.abs(np.swapaxes(B[:256], 1, 2)).astype(np.float64), 1, 2)
        bytes_ = 3 * k * n * n * 4
        res[mode] = {"ms": t, "GB/s": bytes_ / (t * 1e-3) / 1e9, "hbm_frac": bytes_ / (t * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "TFLOP/s": 2 * k * n ** 3 / (t * 1e-3) / 1e12,
                     "max_err_rel_AB": float((err / absAB).max())}
        del ref
    return {"metric": "batched GEMM 64x64x64 x16384 (fp32 / bf16x3)", "workload": "gemm", "results": res}


def bench_cholesky_small(a):
    import torch
    from gputils_amd import api, workloads
    out = {}
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
        t = statistics.median(_events_time_ms(fn, a.steps, refresh))
        bytes_ = k * (32 * 33 // 2 * 8 * 2 + 2 * 32 * 8)
        out[str(k)] = {"ms": t, "solves_per_s": k / (t * 1e-3), "GB/s": bytes_ / (t * 1e-3) / 1e9}
    return {"metric": "FP64 batched Cholesky factor+solve 32x32", "workload": "cholesky_small", "results": out}


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
    fn = {"cholesky_mixed": bench_cholesky_mixed, "gemm": bench_gemm,
          "cholesky_small": bench_cholesky_small}[a.workload]
    out = fn(a)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
