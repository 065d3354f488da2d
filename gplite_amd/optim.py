"""Hyperparameter optimisation on the batched form.

gp_optim (R/gp_optim.R:63-125) refits the model once per energy evaluation.
Its convergence check (R/gp_optim.R:127-160) refits 2P perturbed models, and
restarts repeat the search.  The evaluations of one round are independent,
so here one batched device pass (gpl_batch_energy_grad) evaluates all of them:

* check_convergence_batched: all 2P probes of R/gp_optim.R:143-147 in one pass;
* gp_optim_multistart: K L-BFGS searches in lock step. Each iteration, and each
  backtracking trial, of every search is one problem of the same pass.

Both cover the Gaussian likelihood, as the batched form does.
"""

from __future__ import annotations

import numpy as np

from . import api
from .batch import GPBatch


def _free_names(gp):
    return list(api.get_param(gp).keys())


def check_convergence_batched(gp, x, y, tol_param=1e-1, batch=None):
    """Energies at p_opt +/- tol_param e_k for every log-parameter k, one pass.

    Returns (e_opt, e_minus[P], e_plus[P], converged).  converged follows the
    reference's rule: no probe lowers the energy by more than 1e-2
    (R/gp_optim.R:150-157).  `batch` may pass a GPBatch of 2P copies to reuse.
    """
    p0 = np.array(list(api.get_param(gp).values()), dtype=float)
    P = len(p0)
    params = []
    for k in range(P):
        for sgn in (-1.0, 1.0):
            p = p0.copy()
            p[k] += sgn * tol_param
            params.append(p)
    if batch is None:
        batch = GPBatch([gp] * (2 * P), [x] * (2 * P), [y] * (2 * P), refine_iters=4)
    batch.set_params(params)
    e, _ = batch.energy_grad()
    g = api.gp_fit(gp, x, y)
    e_opt = api.gp_energy(g)
    e_minus, e_plus = e[0::2], e[1::2]
    converged = bool(e_opt - min(e_minus.min(), e_plus.min()) <= 1e-2)
    return e_opt, e_minus, e_plus, converged


def gp_optim_multistart(gp, x, y, starts=8, spread=1.0, maxiter=200, gtol=1e-5, ftol=1e-10, memory=10, seed=0,
                        refine_iters=4, polish=20, return_info=False):
    """Minimise the energy (R/gp.R:138-147) from `starts` initial log-parameter
    vectors (the model's own and starts-1 draws p0 + spread N(0, I)) with
    lock-step L-BFGS.

    Every energy+gradient evaluation of an iteration comes from one batched
    device pass.  Each search uses the two-loop recursion with `memory` pairs
    and Armijo backtracking (c1 = 1e-4, halving).  A search stops when
    max |grad| / n < gtol or its energy change is below ftol·(1+|E|).

    The batched pass computes in mixed precision, whose energies carry
    rounding of about 1e-7 relative: near an optimum the ftol test then stops
    a search wherever that rounding hides the decrease.  The best final point
    is therefore polished by up to `polish` FP64 L-BFGS-B iterations on single
    device fits (gradient tolerance gtol on max |grad| / n); polish=0 skips it.

    Returns the GP refitted (FP64) at the final point, and optionally an info
    dict (energies and iteration counts per start, number of passes, FP64
    polish iterations).
    """
    x = api._as_matrix(x)
    y = np.asarray(y, dtype=float).ravel()
    n = len(y)
    if not isinstance(gp.lik, api.lik_gaussian):
        raise NotImplementedError("the batched form covers the Gaussian likelihood")
    p0 = np.array(list(api.get_param(gp).values()), dtype=float)
    P = len(p0)
    rng = np.random.default_rng(seed)
    X0 = np.vstack([p0] + [p0 + spread * rng.standard_normal(P) for _ in range(starts - 1)])
    batch = GPBatch([gp] * starts, [x] * starts, [y] * starts, refine_iters=refine_iters)
    passes = 0

    def evaluate(Pm):
        nonlocal passes
        batch.set_params(list(Pm))
        e, g = batch.energy_grad(allow_not_pd=True)  # a failed factor is energy +Inf: the step is rejected
        passes += 1
        return e / n, g / n

    p = X0.copy()
    f, g = evaluate(p)
    S = [[] for _ in range(starts)]
    Yh = [[] for _ in range(starts)]
    active = np.isfinite(f)  # a start whose factor fails is dropped
    iters = np.zeros(starts, dtype=int)
    for _ in range(maxiter):
        if not active.any():
            break
        d = np.zeros_like(p)
        for k in np.flatnonzero(active):
            q = g[k].copy()
            alph = []
            for s_, y_ in zip(reversed(S[k]), reversed(Yh[k])):
                a = (s_ @ q) / (y_ @ s_)
                alph.append(a)
                q -= a * y_
            if S[k]:
                q *= (S[k][-1] @ Yh[k][-1]) / (Yh[k][-1] @ Yh[k][-1])
            else:
                q *= min(1.0, 1.0 / max(np.max(np.abs(g[k])), 1e-12))
            for (s_, y_), a in zip(zip(S[k], Yh[k]), reversed(alph)):
                b = (y_ @ q) / (y_ @ s_)
                q += s_ * (a - b)
            d[k] = -q
            if g[k] @ d[k] >= 0:  # not a descent direction: reset the memory
                S[k].clear()
                Yh[k].clear()
                d[k] = -g[k] * min(1.0, 1.0 / max(np.max(np.abs(g[k])), 1e-12))
        t = np.where(active, 1.0, 0.0)
        pending = active.copy()
        p_new, f_new, g_new = p.copy(), f.copy(), g.copy()
        for _ls in range(30):
            trial = p + t[:, None] * d
            ft, gt = evaluate(np.where(pending[:, None], trial, p_new))
            ok = pending & (ft <= f + 1e-4 * t * np.einsum("ij,ij->i", g, d)) & np.isfinite(ft)
            p_new[ok], f_new[ok], g_new[ok] = trial[ok], ft[ok], gt[ok]
            pending &= ~ok
            if not pending.any():
                break
            t = np.where(pending, 0.5 * t, t)
        stalled = pending  # no acceptable step: stop those searches
        for k in np.flatnonzero(active & ~stalled):
            s_, y_ = p_new[k] - p[k], g_new[k] - g[k]
            if s_ @ y_ > 1e-12 * np.linalg.norm(s_) * np.linalg.norm(y_):
                S[k].append(s_)
                Yh[k].append(y_)
                if len(S[k]) > memory:
                    S[k].pop(0)
                    Yh[k].pop(0)
            iters[k] += 1
        df = np.abs(f_new - f)
        p, f, g = p_new, f_new, g_new
        done = (np.max(np.abs(g), axis=1) < gtol) | (df < ftol * (1 + np.abs(f))) | stalled
        active &= ~done
    best = int(np.argmin(f))
    p_final, polished = p[best], 0
    if polish > 0:
        from scipy.optimize import minimize

        def fg(q):
            gq = api.gp_fit(api.set_param(gp, q), x, y, want_grad=True)
            return api.gp_energy(gq) / n, api.gp_energy_grad(gq) / n

        res = minimize(fg, p_final, jac=True, method="L-BFGS-B",
                       options={"maxiter": polish, "gtol": gtol, "ftol": 1e-15})
        if np.all(np.isfinite(res.x)) and np.isfinite(res.fun):
            p_final, polished = res.x, int(res.nit)
    g_best = api.gp_fit(api.set_param(gp, p_final), x, y)
    if return_info:
        return g_best, {"energy_per_n": f, "params": p, "iterations": iters, "passes": passes, "best": best,
                        "polish_iterations": polished, "names": _free_names(gp)}
    return g_best
