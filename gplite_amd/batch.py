"""Batched form: energy and gradient for many independent Gaussian-likelihood
GPs (hyperparameter restarts, LOO folds, independent datasets) in one device
pass (gpl_batch_* in include/gplite_amd.h)."""

from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import _lib as L
from . import api


class GPBatch:
    """gps: list of GP objects with the same covariance structure and a
    Gaussian likelihood; xs / ys: per-problem inputs and targets."""

    def __init__(self, gps, xs, ys, refine_iters=4, jitter=1e-6):
        if not gps:
            raise ValueError("empty batch")
        self.gps = [g.copy() for g in gps]
        for g in self.gps:
            if not isinstance(g.lik, api.lik_gaussian):
                raise NotImplementedError("the batched form covers the Gaussian likelihood")
        self.xs = [api._as_matrix(x) for x in xs]
        self.ys = [np.ascontiguousarray(np.asarray(y, dtype=np.float64).ravel()) for y in ys]
        for g, x in zip(self.gps, self.xs):
            for c in g.cfs:
                c.learn_scales(x)
        terms, xps = self._terms()
        self.nterms = terms[0][1]
        self._xps = xps
        n = (C.c_int64 * len(xps))(*[x.shape[0] for x in xps])
        nc = (C.c_int64 * len(xps))(*[x.shape[1] for x in xps])
        xptr = (L._dp * len(xps))(*[L.dptr(x) for x in xps])
        yptr = (L._dp * len(xps))(*[L.dptr(y) for y in self.ys])
        flat = self._flat(terms)
        sig = np.array([g.lik.sigma for g in self.gps])
        o = L.Opts()
        L.lib().gpl_opts_default(C.byref(o))
        o.refine_iters = refine_iters
        o.jitter = jitter
        h = L._vp()
        L.check(L.lib().gpl_batch_create(L.ctx(), len(xps), flat, self.nterms, xptr, n, nc, yptr, L.dptr(sig),
                                         C.byref(o), C.byref(h)))
        self.handle = h
        self.ng = L.lib().gpl_batch_grad_len(h)
        _, _, _, mask = api._program(self.gps[0].cfs, self.xs[0])
        mask = mask + [0 if api._fixed(self.gps[0].lik.priors.get("sigma")) else 1]
        self._mask = np.asarray(mask, dtype=bool)

    def _terms(self):
        out, xps = [], []
        for g, x in zip(self.gps, self.xs):
            t, nt, xp, _ = api._program(g.cfs, x)
            out.append((t, nt))
            xps.append(xp)
        return out, xps

    def _flat(self, terms):
        nt = terms[0][1]
        arr = (L.CfTerm * (nt * len(terms)))()
        for p, (t, ntp) in enumerate(terms):
            if ntp != nt:
                raise ValueError("all problems need the same number of covariance terms")
            for k in range(nt):
                arr[p * nt + k] = t[k]
        return arr

    def set_params(self, params):
        """params: list of log-parameter vectors (get_param order), one per problem."""
        self.gps = [api.set_param(g, p) for g, p in zip(self.gps, params)]
        terms, _ = self._terms()
        sig = np.array([g.lik.sigma for g in self.gps])
        L.check(L.lib().gpl_batch_set_params(self.handle, self._flat(terms), L.dptr(sig)))

    def energy_grad(self, include_prior=True, allow_not_pd=False):
        """(energies [nprob], gradients [nprob, P]) in get_param order.  With
        allow_not_pd a problem whose factor fails, or whose refinement does not
        converge, gets energy +Inf and NaN gradients instead of raising
        (status() tells which)."""
        np_ = len(self.gps)
        lz = np.empty(np_)
        g = np.empty(np_ * self.ng)
        rc = L.lib().gpl_batch_energy_grad(self.handle, L.dptr(lz), L.dptr(g))
        L.check(rc, allow=(L.GPL_ERR_NOT_PD, L.GPL_ERR_NOCONV) if allow_not_pd else ())
        g = g.reshape(np_, self.ng)[:, self._mask]
        e = -lz
        if include_prior:
            e = e - np.array([api.lpdf_prior(gp) for gp in self.gps])
            g = g - np.array([api._lpdf_prior_grad(gp) for gp in self.gps])
        return e, g

    def status(self):
        """Per-problem status of the last pass (GPL_OK, GPL_ERR_NOT_PD, GPL_ERR_NOCONV)
        and the last refinement correction relative to each problem's solution."""
        np_ = len(self.gps)
        st = np.empty(np_, dtype=np.int32)
        rel = np.empty(np_)
        L.check(L.lib().gpl_batch_status(self.handle, st.ctypes.data_as(C.POINTER(C.c_int32)), L.dptr(rel)))
        return st, rel

    def info(self):
        nb = C.c_int32()
        rel = C.c_double()
        L.check(L.lib().gpl_batch_info(self.handle, C.byref(nb), C.byref(rel)))
        return {"buckets": nb.value, "last_refine_rel": rel.value}

    def tune(self, path=None):
        """Time tile edge x panel width per bucket; optionally persist the table as JSON."""
        cap = 256
        bn = (C.c_int32 * cap)(); bm = (C.c_int32 * cap)(); nb = (C.c_int32 * cap)(); ms = (C.c_double * cap)()
        cnt = C.c_int32()
        L.check(L.lib().gpl_batch_tune(self.handle, cap, bn, bm, nb, ms, C.byref(cnt)))
        k = cnt.value
        table = [{"n": bn[i], "tile": bm[i], "panel": nb[i], "ms": ms[i]} for i in range(min(k, cap))]
        if path:
            with open(path, "w") as f:
                json.dump(table, f, indent=1)
        return table

    def load_tuning(self, table):
        if isinstance(table, str):
            with open(table) as f:
                table = json.load(f)
        k = len(table)
        bn = (C.c_int32 * k)(*[t["n"] for t in table])
        bm = (C.c_int32 * k)(*[t["tile"] for t in table])
        nb = (C.c_int32 * k)(*[t["panel"] for t in table])
        L.check(L.lib().gpl_batch_set_tuning(self.handle, k, bn, bm, nb))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and h.value and L._lib is not None:
            L._lib.gpl_batch_free(h)
            self.handle = None


def gp_fit_batch(gps, xs, ys, refine_iters=6, jitter=1e-6):
    """gp_fit for many Gaussian-likelihood GPs of the same covariance structure in
    one batched device pass (gpl_fit_batch): returns fitted GP objects whose
    resident mixed-precision models predict, draw and differentiate like those of
    gp_fit(..., precision="mixed").  A problem whose FP32 factor or refinement
    fails is fitted again on its own by gp_fit (which redoes it in FP64)."""
    gps = [g.copy() for g in gps]
    for g in gps:
        if not isinstance(g.lik, api.lik_gaussian):
            raise NotImplementedError("the batched fit covers the Gaussian likelihood")
    xs = [api._as_matrix(x) for x in xs]
    ys = [np.ascontiguousarray(np.asarray(y, dtype=np.float64).ravel()) for y in ys]
    for g, x in zip(gps, xs):
        for c in g.cfs:
            c.learn_scales(x)
    progs = [api._program(g.cfs, x) for g, x in zip(gps, xs)]
    nt = progs[0][1]
    flat = (L.CfTerm * (nt * len(gps)))()
    for p, (t, ntp, _, _) in enumerate(progs):
        if ntp != nt:
            raise ValueError("all problems need the same number of covariance terms")
        for k in range(nt):
            flat[p * nt + k] = t[k]
    xps = [pr[2] for pr in progs]
    npb = len(gps)
    n = (C.c_int64 * npb)(*[x.shape[0] for x in xps])
    nc = (C.c_int64 * npb)(*[x.shape[1] for x in xps])
    xptr = (L._dp * npb)(*[L.dptr(x) for x in xps])
    yptr = (L._dp * npb)(*[L.dptr(y) for y in ys])
    sig = np.array([g.lik.sigma for g in gps])
    o = L.Opts()
    L.lib().gpl_opts_default(C.byref(o))
    o.refine_iters = refine_iters
    o.jitter = jitter
    o.precision = L.PREC_MIXED
    handles = (L._vp * npb)()
    lz = np.empty(npb)
    st = np.empty(npb, dtype=np.int32)
    L.check(L.lib().gpl_fit_batch(L.ctx(), npb, flat, nt, xptr, n, nc, yptr, L.dptr(sig), C.byref(o), handles,
                                  L.dptr(lz), st.ctypes.data_as(C.POINTER(C.c_int32))))
    out = []
    for p, g in enumerate(gps):
        if st[p] != L.GPL_OK:
            out.append(api.gp_fit(g, xs[p], ys[p], jitter=jitter, precision="mixed"))
            continue
        g.fit = api._Fit(L._vp(handles[p]), ys[p].shape[0], float(lz[p]), 1, L.APPROX_LAPLACE)
        g.x = xs[p]
        g.fitted = True
        g.precision = "mixed"
        g._trials, g._offset = None, None
        out.append(g)
    return out
