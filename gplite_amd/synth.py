"""Synthetic workloads with fixed seeds (numpy PCG64).  Every path (device,
oracle, CPU baseline) is fed the same arrays.

c1: 1-D regression, n=200, cf_sexp(0.3, 1), lik_gaussian(0.1), pred at 300 points in [-4, 4]
c2: n=8192, d=8, x~U[0,1]^8, Matern-5/2 with ARD lscale~U[0.2,1], magn=1, y = GP sample + 0.05 N(0,1)
c3: n=4096, d=4, x~N(0,I), f~GP(sexp(0.8, 2)), y~Bernoulli(logistic(f))
"""

from __future__ import annotations

import numpy as np


def _gp_sample(K, rng, jitter=1e-8):
    import scipy.linalg as sla
    n = K.shape[0]
    K = K.copy()
    K[np.diag_indices(n)] += jitter
    Lc = sla.cholesky(K, lower=True, overwrite_a=True, check_finite=False)
    return Lc @ rng.standard_normal(n)


def _matern52_ard(x1, x2, ls, magn=1.0, chunk=1024):
    out = np.empty((x1.shape[0], x2.shape[0]))
    w = 1.0 / (ls * ls)
    for s in range(0, x1.shape[0], chunk):
        d = x1[s:s + chunk, None, :] - x2[None, :, :]
        r = np.sqrt(np.einsum("ijk,ijk,k->ij", d, d, w))
        out[s:s + chunk] = magn * magn * (1 + np.sqrt(5) * r + 5 * r * r / 3) * np.exp(-np.sqrt(5) * r)
    return out


def _sexp(x1, x2, ls, magn, chunk=1024):
    out = np.empty((x1.shape[0], x2.shape[0]))
    for s in range(0, x1.shape[0], chunk):
        d = x1[s:s + chunk, None, :] - x2[None, :, :]
        out[s:s + chunk] = magn * magn * np.exp(-0.5 * np.einsum("ijk,ijk->ij", d, d) / ls ** 2)
    return out


def c1(seed=32004):
    r = np.random.default_rng(seed)
    n = 200
    x = r.standard_normal(n)
    y = np.sin(3 * x) * np.exp(-np.abs(x)) + 0.1 * r.standard_normal(n)
    xt = np.linspace(-4, 4, 300)
    return dict(x=x[:, None], y=y, xt=xt[:, None], lscale=0.3, magn=1.0, sigma=0.1)


def c2(n=8192, d=8, m=4096, seed=82008):
    r = np.random.default_rng(seed)
    x = r.uniform(0, 1, (n, d))
    ls = r.uniform(0.2, 1.0, d)
    K = _matern52_ard(x, x, ls)
    f = _gp_sample(K, r)
    del K
    y = f + 0.05 * r.standard_normal(n)
    xt = r.uniform(0, 1, (m, d))
    return dict(x=x, y=y, xt=xt, lscale=ls, magn=1.0, sigma=0.05)


def c3(n=4096, d=4, seed=43004):
    r = np.random.default_rng(seed)
    x = r.standard_normal((n, d))
    K = _sexp(x, x, 0.8, 2.0)
    f = _gp_sample(K, r)
    del K
    y = (r.uniform(size=n) < 1 / (1 + np.exp(-f))).astype(float)
    return dict(x=x, y=y, lscale=0.8, magn=2.0)


def c4(nprob=2048, nmin=64, nmax=2048, seed=44096):
    """nprob independent regressions: n ~ round(exp(U[ln nmin, ln nmax])), d ~ U{1..8},
    x ~ U[-2,2]^d, lscale ~ U[0.3, 1.5], magn = 1, y = sin(sum x) + 0.1 N(0,1), sigma = 0.1."""
    r = np.random.default_rng(seed)
    out = []
    for _ in range(nprob):
        n = int(round(np.exp(r.uniform(np.log(nmin), np.log(nmax)))))
        d = int(r.integers(1, 9))
        x = r.uniform(-2, 2, (n, d))
        y = np.sin(x.sum(1)) + 0.1 * r.standard_normal(n)
        out.append(dict(x=x, y=y, lscale=float(r.uniform(0.3, 1.5)), magn=1.0, sigma=0.1))
    return out
