"""Benchmark: GP energy + hyperparameter-gradient evaluations per second on one
MI355X.  Workload: n=8192, d=8, x~U[0,1]^8, Matern-5/2 with ARD length-scales,
Gaussian likelihood (sigma=0.05), y a seeded GP sample.  One step = one
gp_fit (covariance build, B = I + K/sigma^2 Cholesky, alpha, log evidence) plus
one gp_energy_grad (B^-1 and the fused gradient reduction), inputs resident on
the GPU.

Both device forms are timed: plain FP64 and the mixed-precision form (FP32
MFMA, FP64 diagonal blocks / log-det / refinement).  `value` is the mixed
form's rate; its log-evidence, gradient and predictive-mean errors against the
plain FP64 path are reported next to it.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 8192] [--no-cpu]
Prints one JSON line on stdout.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6, "f32": 157.3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--forms", default="fp64,mixed,mixed_batched")
    ap.add_argument("--restarts", type=int, default=16, help="problems per pass in the batched form")
    ap.add_argument("--refine", type=int, default=3, help="refinement sweeps in the batched form")
    ap.add_argument("--tuning", default=None,
                    help="tile,panel for the batched form (skips the autotuner), e.g. 128,512")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; how the path would shard is described in "
                                   "DESIGN.md but not built", "n_gpus": args.gpus}))
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
    _, _, xtp, _ = api._program(gp.cfs, w["xt"])
    y = np.ascontiguousarray(w["y"])
    lk = api._lik_struct(gp.lik)
    ng = lib.gpl_grad_len(terms, nt, L.LIK_GAUSSIAN)

    def run_form(prec):
        o = L.Opts()
        lib.gpl_opts_default(C.byref(o))
        o.flags = 1  # GPL_FLAG_INVERSE: the gradient follows every fit
        o.precision = L.PREC_MIXED if prec == "mixed" else L.PREC_FP64
        handle = C.c_void_p()
        logz = C.c_double()
        grad = np.empty(ng)

        def step():
            L.check(lib.gpl_fit(ctx, terms, nt, L.dptr(xp), n, xp.shape[1], L.dptr(y), None, None, C.byref(lk),
                                C.byref(o), C.byref(handle), C.byref(logz)))
            L.check(lib.gpl_energy_grad(handle, L.dptr(grad)))

        for i in range(args.warmup):
            step()
        lat = []
        t_all = time.perf_counter()
        for i in range(args.steps):
            ts = time.perf_counter()
            step()  # each call ends with a stream synchronisation
            lat.append(time.perf_counter() - ts)
        elapsed = time.perf_counter() - t_all
        log(f"[{prec}] {args.steps} steps in {elapsed:.3f}s -> {args.steps / elapsed:.2f} evals/s")
        # device time per kernel class: HIP events around each launch group (separate pass)
        L.profile(enable=True, reset=True)
        step()
        prof = L.profile(enable=False)
        # prediction at the 4096 test points (mean and variance), once
        m = xtp.shape[0]
        mean = np.empty(m); var = np.empty(m)
        tp = time.perf_counter()
        L.check(lib.gpl_pred(handle, L.dptr(xtp), m, 1, L.dptr(mean), L.dptr(var)))
        tpred = time.perf_counter() - tp
        lat_ms = np.array(lat) * 1e3
        res = {"evals_per_s": args.steps / elapsed, "ms_per_step": 1e3 * elapsed / args.steps,
               "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
               "logz": logz.value, "grad": grad.copy(), "pred_mean": mean, "pred_var": var,
               "pred_ms": tpred * 1e3, "prof": prof}
        lib.gpl_model_free(handle)
        return res

    def run_batched(nrest):
        """nrest hyperparameter restarts of the same regression (ARD length-scales
        perturbed by exp(0.2 N(0,1)); restart 0 keeps the base values) in one
        batched mixed-precision pass per step."""
        from gplite_amd.batch import GPBatch
        rr = np.random.default_rng(7)
        gps = []
        for i in range(nrest):
            ls = w["lscale"] * (np.exp(0.2 * rr.standard_normal(d)) if i else 1.0)
            gps.append(G.gp_init(G.cf_matern52(lscale=ls, magn=w["magn"]), G.lik_gaussian(sigma=w["sigma"])))
        Bt = GPBatch(gps, [w["x"]] * nrest, [y] * nrest, refine_iters=args.refine)
        if args.tuning:
            tile, panel = (int(v) for v in args.tuning.split(","))
            table = [{"n": -(-n // 64) * 64, "tile": tile, "panel": panel}]
            Bt.load_tuning(table)
        else:
            table = Bt.tune()
        log(f"[batched] tuning: {table}")
        ngb = lib.gpl_batch_grad_len(Bt.handle)
        lz = np.empty(nrest)
        gb = np.empty(nrest * ngb)

        def step():
            L.check(lib.gpl_batch_energy_grad(Bt.handle, L.dptr(lz), L.dptr(gb)))

        for i in range(args.warmup):
            step()
        lat = []
        t_all = time.perf_counter()
        for i in range(args.steps):
            ts = time.perf_counter()
            step()
            lat.append(time.perf_counter() - ts)
        elapsed = time.perf_counter() - t_all
        log(f"[mixed_batched] {args.steps} passes x {nrest} problems in {elapsed:.3f}s -> "
            f"{args.steps * nrest / elapsed:.2f} evals/s; refine rel {Bt.info()['last_refine_rel']:.2e}")
        L.profile(enable=True, reset=True)
        step()
        prof = L.profile(enable=False)
        lat_ms = np.array(lat) * 1e3
        res = {"evals_per_s": args.steps * nrest / elapsed, "ms_per_step": 1e3 * elapsed / args.steps,
               "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
               "logz": float(lz[0]), "grad": gb[:ngb].copy(), "prof": prof, "problems": nrest,
               "tuning": table, "refine_rel": Bt.info()["last_refine_rel"]}
        del Bt
        return res

    forms = args.forms.split(",")
    R = {f: (run_batched(args.restarts) if f == "mixed_batched" else run_form(f)) for f in forms}
    head = "mixed_batched" if "mixed_batched" in R else ("mixed" if "mixed" in R else forms[0])
    H = R[head]
    prof = H["prof"]
    # roofline of the B^-1 = X^T X product: the largest single kernel launch of a
    # step (N^3/3 flops per problem, one launch per pass), timed with HIP events
    # on the stream it runs on during a separate profiled pass
    rk = {"mixed_batched": "batch_lauum_f32", "mixed": "lauum_f32", "fp64": "lauum"}[head]
    rc = prof[rk]
    ach = rc["flops"] / (rc["ms"] * 1e-3) / 1e12
    pk = PEAK["f32"] if rk.endswith("f32") else PEAK["f64"]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            tt = json.load(fh).get(rk)
        if tt and tt.get("n") == n and tt.get("problems", 1) == H.get("problems", 1):
            traffic = tt["bytes_per_launch"]
    roof = {"bound": "mfma", "kernel": rk, "achieved": round(ach, 3), "peak": pk, "unit": "TFLOP/s",
            "frac": round(ach / pk, 4), "traffic": traffic, "launches": rc["launches"],
            "avg_launch_us": round(rc["ms"] * 1e3 / max(rc["launches"], 1), 2),
            "algorithmic_flops_per_launch": rc["flops"] / max(rc["launches"], 1),
            "top_classes": {k: {"ms": round(v["ms"], 3), "TFLOP/s": round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 1)}
                            for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])[:4]}}
    for f, r in R.items():
        log(f"[{f}] kernel classes:", json.dumps({k: {"ms": round(v["ms"], 3), "n": v["launches"],
                                                       "TF": round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 1)}
                                                   for k, v in sorted(r["prof"].items(), key=lambda kv: -kv[1]["ms"])}))
    parity = None
    if "fp64" in R and "mixed" in R:
        a, b = R["mixed"], R["fp64"]
        parity = {"rel_logz_mixed_vs_fp64": abs(a["logz"] / b["logz"] - 1),
                  "rel_grad_mixed_vs_fp64": float(np.max(np.abs(a["grad"] - b["grad"])) / np.max(np.abs(b["grad"]))),
                  "rel_pred_mean_mixed_vs_fp64": float(np.max(np.abs(a["pred_mean"] - b["pred_mean"]))
                                                       / np.max(np.abs(b["pred_mean"]))),
                  "rel_pred_var_mixed_vs_fp64": float(np.max(np.abs(a["pred_var"] - b["pred_var"]))
                                                      / np.max(np.abs(b["pred_var"])))}
        if "mixed_batched" in R:
            c = R["mixed_batched"]
            parity["rel_logz_batched_vs_fp64"] = abs(c["logz"] / b["logz"] - 1)
            parity["rel_grad_batched_vs_fp64"] = float(np.max(np.abs(c["grad"] - b["grad"])) / np.max(np.abs(b["grad"])))
        log("parity:", parity)

    cpu = None
    if not args.no_cpu:
        from oracle.cpu_baseline import energy_grad_matern52_ard
        try:
            from threadpoolctl import threadpool_info
            cores = max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
        except Exception:
            cores = os.cpu_count()
        lz, gcpu, secs = energy_grad_matern52_ard(w["x"], y, w["lscale"], w["magn"], w["sigma"])
        ref = R.get("fp64", H)
        cpu = {"value": round(1.0 / secs, 5), "unit": "evals/s", "cores": cores, "kind": "port",
               "sample": f"one energy+gradient evaluation at n={n}, d={d}: numpy/scipy LAPACK dpotrf + dpotri "
                         f"restatement of the reference sequence ({secs:.1f}s)",
               "rel_logz_fp64_gpu_vs_cpu": abs(ref["logz"] / lz - 1),
               "rel_grad_fp64_gpu_vs_cpu": float(np.max(np.abs(ref["grad"] - gcpu)) / np.max(np.abs(gcpu)))}
        log("cpu baseline:", cpu)

    out = {
        "metric": "GP energy+grad evals/s (n=8192, d=8, Matern-5/2 ARD) & p50/p99 latency vs plain FP64, rel logZ err",
        "value": round(H["evals_per_s"], 4),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(H["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head.startswith("mixed") else "f64",
        "data": "synthetic (seeded Matern-5/2 GP sample, numpy PCG64)",
        "config": {"workload": "energy+gradient per step, n=8192, d=8, cf_matern52 ARD, lik_gaussian(sigma=0.05); "
                               "pred at 4096 points; headline form: " + head
                               + (f" ({args.restarts} hyperparameter restarts per pass)" if head == "mixed_batched" else ""),
                    "n": n, "d": d, "form": head},
        "latency_ms": {"p50": round(H["p50_ms"], 3), "p99": round(H["p99_ms"], 3)},
        "forms": {f: {"evals_per_s": round(r["evals_per_s"], 3), "ms_per_step": round(r["ms_per_step"], 3),
                      "p50_ms": round(r["p50_ms"], 3), "p99_ms": round(r["p99_ms"], 3),
                      **({"pred_4096_ms": round(r["pred_ms"], 3)} if "pred_ms" in r else {}),
                      **({"problems_per_pass": r["problems"]} if "problems" in r else {})} for f, r in R.items()},
        "speedup_vs_fp64": ({f: round(r["evals_per_s"] / R["fp64"]["evals_per_s"], 3) for f, r in R.items()
                             if f != "fp64"} if "fp64" in R else None),
        "parity": parity,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
