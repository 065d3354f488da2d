"""Serving form: continuous batching of fit / predict requests on one GPU
(gpl_server_* in include/gplite_amd.h), plus a Poisson load generator that
reports p50 / p99 latency and throughput."""

from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib as L
from . import api


class Server:
    def __init__(self, workers=4, mem_gb=64.0, device=0):
        h = L._vp()
        L.lib()
        L.check(L.lib().gpl_server_create(device, workers, int(mem_gb * 2 ** 30), C.byref(h)))
        self.h = h
        self._keep = {}

    def submit_fit(self, gp, x, y, trials=None, offset=None, precision="fp64", approx_tol=None):
        """precision: "fp64", "mixed", or "auto": mixed for a Gaussian likelihood
        (one factorisation, where FP32 updates with FP64 refinement are faster) and
        FP64 for Laplace and EP fits, whose every step would also form L^-1 for the
        refinement (n = 4096 Laplace: 31.7 fits/s in FP64 against 22.7 mixed)."""
        g = gp.copy()
        if precision == "auto":
            precision = "mixed" if isinstance(g.lik, api.lik_gaussian) else "fp64"
        x = api._as_matrix(x)
        y = np.ascontiguousarray(np.asarray(y, dtype=np.float64).ravel())
        for c in g.cfs:
            c.learn_scales(x)
        terms, nt, xp, _ = api._program(g.cfs, x)
        o = L.Opts()
        L.lib().gpl_opts_default(C.byref(o))
        if isinstance(g.approx, api.approx_ep):
            o.approx, o.maxiter, o.damping, o.quad_order = L.APPROX_EP, g.approx.maxiter, g.approx.damping, g.approx.quad_order
        else:
            o.maxiter, o.tol = g.approx.maxiter, g.approx.tol
        o.precision = L.PREC_MIXED if precision == "mixed" else L.PREC_FP64
        lk = api._lik_struct(g.lik)
        tr = None if trials is None else np.ascontiguousarray(np.asarray(trials, float))
        off = None if offset is None else np.ascontiguousarray(np.broadcast_to(np.asarray(offset, float), y.shape))
        mid = C.c_int64(); tk = C.c_int64()
        L.check(L.lib().gpl_server_submit_fit(self.h, terms, nt, L.dptr(xp), xp.shape[0], xp.shape[1], L.dptr(y),
                                              L.dptr(tr), L.dptr(off), C.byref(lk), C.byref(o), C.byref(mid),
                                              C.byref(tk)))
        self._keep[("model", mid.value)] = g  # cfs with learned input scales, for prediction inputs
        return mid.value, tk.value

    def submit_pred(self, model_id, xt, var=True):
        g = self._keep[("model", model_id)]
        _, _, xtp, _ = api._program(g.cfs, api._as_matrix(xt))
        m = xtp.shape[0]
        mean = np.empty(m)
        v = np.empty(m) if var else None
        tk = C.c_int64()
        L.check(L.lib().gpl_server_submit_pred(self.h, model_id, L.dptr(xtp), m, int(bool(var)), L.dptr(mean),
                                               L.dptr(v), C.byref(tk)))
        self._keep[("pred", tk.value)] = (xtp, mean, v)
        return tk.value

    def wait(self, ticket):
        lat = C.c_double(); q = C.c_double(); lz = C.c_double(); co = C.c_int32()
        rc = L.lib().gpl_server_wait(self.h, ticket, C.byref(lat), C.byref(q), C.byref(lz), C.byref(co))
        L.check(rc, allow=(L.GPL_ERR_NOCONV,))
        out = {"latency_ms": lat.value, "queue_ms": q.value, "log_evidence": lz.value, "coalesced": co.value}
        kept = self._keep.pop(("pred", ticket), None)
        if kept is not None:
            out["mean"], out["var"] = kept[1], kept[2]
        return out

    def release(self, model_id):
        L.check(L.lib().gpl_server_release_model(self.h, model_id))
        self._keep.pop(("model", model_id), None)

    def stats(self):
        v = [C.c_int64() for _ in range(6)]
        L.check(L.lib().gpl_server_stats(self.h, *[C.byref(x) for x in v]))
        return dict(zip(["fits", "preds", "launches", "evictions", "bytes_in_use", "resident_models"],
                        [x.value for x in v]))

    def close(self):
        if self.h and self.h.value:
            L.lib().gpl_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def c5_fit_request(rng, nmin=128, nmax=16384, d=6):
    """One serving fit request: n ~ logU[nmin, nmax], x ~ U[0,1]^6, cf_matern32(lscale ~ U[0.2,0.8]);
    Gaussian (sigma=0.1) or Poisson counts with a Laplace fit, 50/50."""
    n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
    x = rng.uniform(0, 1, (n, d))
    f = np.sin(2.0 * x.sum(1)) + 0.5 * np.cos(3.0 * x[:, 0])
    cf = api.cf_matern32(lscale=float(rng.uniform(0.2, 0.8)), magn=1.0)
    if rng.uniform() < 0.5:
        gp = api.gp_init(cf, api.lik_gaussian(sigma=0.1))
        y = f + 0.1 * rng.standard_normal(n)
    else:
        gp = api.gp_init(cf, api.lik_poisson())
        y = rng.poisson(np.exp(f)).astype(float)
    return gp, x, y


def poisson_load(server, rate, nreq, frac_fit=0.3, m=1000, seed=0, nmin=128, nmax=16384, precision="fp64",
                 content_seed=None):
    """Submit nreq requests with exponential inter-arrival times (rate per second);
    returns per-type latency percentiles and throughput.  The arrival times come
    from `seed`; the requests themselves (fit or predict, sizes, data) from
    `content_seed` (default: `seed`), so one content seed serves the same request
    set at every rate and the throughputs of a sweep compare the same work."""
    rng_gap = np.random.default_rng(seed)
    rng = np.random.default_rng(seed if content_seed is None else content_seed)
    gaps = rng_gap.exponential(1.0 / rate, nreq)
    if content_seed is None:
        rng = rng_gap
    kinds = ["fit"] + ["fit" if u < frac_fit else "pred" for u in rng.uniform(size=nreq - 1)]
    models = []
    tickets = []
    fit_n = []
    t0 = time.perf_counter()
    t_next = 0.0
    for k, g in zip(kinds, gaps):
        t_next += g
        now = time.perf_counter() - t0
        if t_next > now:
            time.sleep(t_next - now)
        if k == "fit" or not models:
            gp, x, y = c5_fit_request(rng, nmin, nmax)
            fit_n.append(len(y))
            mid, tk = server.submit_fit(gp, x, y, precision=precision)
            models.append(mid)
            tickets.append(("fit", tk))
        else:
            mid = models[int(rng.integers(0, len(models)))]
            xt = rng.uniform(0, 1, (m, 6))
            tickets.append(("pred", server.submit_pred(mid, xt, var=True)))
    res = {"fit": [], "pred": []}
    co = []
    for k, tk in tickets:
        r = server.wait(tk)
        res[k].append(r["latency_ms"])
        if k == "pred":
            co.append(r["coalesced"])
    t_end = time.perf_counter() - t0
    out = {"rate_req_s": rate, "requests": nreq, "wall_s": t_end, "throughput_req_s": nreq / t_end,
           "offered_s": float(np.sum(gaps)), "mean_pred_coalesced": float(np.mean(co)) if co else None,
           "fit_sum_n3": float(sum(float(n) ** 3 for n in fit_n))}
    allv = res["fit"] + res["pred"]
    for k, v in (("all", allv), ("fit", res["fit"]), ("pred", res["pred"])):
        if v:
            out[k] = {"n": len(v), "p50_ms": float(np.percentile(v, 50)), "p99_ms": float(np.percentile(v, 99))}
    for mid in models:
        server.release(mid)
    return out
