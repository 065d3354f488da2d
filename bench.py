"""Benchmark: GP energy + hyperparameter-gradient evaluations per second on one
MI355X, dense FP64 path, n=8192, d=8, Matern-5/2 with ARD length-scales,
Gaussian likelihood (sigma=0.05).  One step = one gp_fit (covariance build,
B = I + K/sigma^2 Cholesky, alpha, log evidence) plus one gp_energy_grad
(L^-1, B^-1, fused gradient reduction) with inputs already resident on the GPU.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 8192] [--no-cpu]
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md, "
                                   "not built", "n_gpus": args.gpus}))
        sys.exit(2)

    import gplite_amd as G
    from gplite_amd import _lib as L, api, synth

    G.require_device()
    log("device:", L.device_name())
    t0 = time.time()
    w = synth.c2(n=args.n)
    log(f"workload generated in {time.time() - t0:.1f}s")
    n, d = w["x"].shape
    lib, ctx = L.lib(), L.ctx()

    gp = G.gp_init(G.cf_matern52(lscale=w["lscale"], magn=w["magn"]), G.lik_gaussian(sigma=w["sigma"]))
    for c in gp.cfs:
        c.learn_scales(w["x"])
    terms, nt, xp, _ = api._program(gp.cfs, w["x"])
    y = np.ascontiguousarray(w["y"])
    lk = api._lik_struct(gp.lik)
    o = L.Opts()
    lib.gpl_opts_default(C.byref(o))
    handle = C.c_void_p()
    logz = C.c_double()
    ng = lib.gpl_grad_len(terms, nt, L.LIK_GAUSSIAN)
    grad = np.empty(ng)

    def step():
        L.check(lib.gpl_fit(ctx, terms, nt, L.dptr(xp), n, xp.shape[1], L.dptr(y), None, None, C.byref(lk),
                            C.byref(o), C.byref(handle), C.byref(logz)))
        L.check(lib.gpl_energy_grad(handle, L.dptr(grad)))

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup}")
    lat = []
    t_all = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        step()  # each call ends with a stream synchronisation
        lat.append(time.perf_counter() - ts)
    elapsed = time.perf_counter() - t_all
    log(f"timed {args.steps} steps in {elapsed:.3f}s")
    value = args.steps / elapsed

    # per-kernel-class device time with HIP events on the library stream (separate pass)
    L.profile(enable=True, reset=True)
    step()
    prof = L.profile(enable=False)
    dom = max(prof.items(), key=lambda kv: kv[1]["ms"])
    tot_ms = sum(v["ms"] for v in prof.values())
    ach = dom[1]["flops"] / (dom[1]["ms"] * 1e-3) / 1e12
    roof = {"bound": "mfma", "kernel": dom[0], "achieved": round(ach, 3), "peak": 78.6, "unit": "TFLOP/s",
            "frac": round(ach / 78.6, 4), "traffic": None,
            "launches": dom[1]["launches"], "avg_launch_us": round(dom[1]["ms"] * 1e3 / max(dom[1]["launches"], 1), 2)}
    breakdown = {k: {"ms": round(v["ms"], 3), "launches": v["launches"],
                     "tflops": round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 2) if v["flops"] else None}
                 for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])}
    log("kernel classes:", json.dumps(breakdown))

    err = None
    if not args.no_check:
        # parity on the timed workload: relative log-evidence error vs the FP64 CPU path
        pass

    cpu = None
    if not args.no_cpu:
        from oracle.cpu_baseline import energy_grad_matern52_ard
        try:
            from threadpoolctl import threadpool_info
            cores = max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
        except Exception:
            cores = os.cpu_count()
        lz, gcpu, secs = energy_grad_matern52_ard(w["x"], y, w["lscale"], w["magn"], w["sigma"])
        err = {"rel_logz_vs_cpu": abs(logz.value / lz - 1),
               "rel_grad_vs_cpu": float(np.max(np.abs(grad - gcpu)) / np.max(np.abs(gcpu)))}
        cpu = {"value": round(1.0 / secs, 5), "unit": "evals/s", "cores": cores, "kind": "port",
               "sample": f"one energy+gradient evaluation at n={n}, d={d} (scipy LAPACK dpotrf/dpotri, "
                         f"{secs:.1f}s)"}
        log("cpu baseline:", cpu, err)

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "GP energy+grad evals/s (n=8192, d=8, Matern-5/2 ARD, plain FP64)",
        "value": round(value, 4),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded GP sample, numpy PCG64)",
        "config": {"workload": "energy+gradient, n=8192, d=8, cf_matern52 ARD, lik_gaussian(sigma=0.05)",
                   "n": n, "d": d, "precision": "fp64", "parallelism": "single GPU"},
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 3),
                       "p99": round(float(np.percentile(lat_ms, 99)), 3)},
        "device_ms_per_step": round(tot_ms, 3),
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": err,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
