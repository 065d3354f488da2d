"""CPU baseline for the energy+gradient workload (test/benchmark infrastructure).

The reference's CPU path is R + host LAPACK.  R is not available, so the
baseline is this FP64 restatement of the same operation sequence on host
LAPACK through scipy (OpenBLAS threads = the cores given):
K build (src/cf.cpp:86-119 with ARD), B = I + K/sigma^2 and dpotrf
(R/laplace.R:27-28), alpha by two triangular solves (:29), f = K alpha (:30),
log evidence (:33-38), then the gradient the reference lacks: dpotri for
B^-1 and -1/2 sum (a a^T - K_y^-1) o dK/dlog(theta) (Rasmussen & Williams eq. 5.9).
"""

from __future__ import annotations

import math
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.linalg as sla
from scipy.linalg import lapack


def host_threads():
    """(nproc, threads): the CPUs this process may run on, and the thread count the
    baseline uses -- all of them, unless OMP_NUM_THREADS names the host's share
    (a machine that gives each GPU a fixed share of its cores sets it)."""
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(nproc, int(env)) if env.isdigit() and int(env) > 0 else nproc
    return nproc, threads


def _threads():
    """Worker count for the elementwise passes: the BLAS thread count in force
    (so threadpool_limits(1) gives a single-thread run), else the CPU count."""
    try:
        from threadpoolctl import threadpool_info
        return max(1, max((i.get("num_threads", 1) for i in threadpool_info()), default=1))
    except Exception:
        return os.cpu_count() or 1


def _m52_parts(x1, x2, w, magn):
    d = x1[:, None, :] - x2[None, :, :]
    t2 = d * d * w
    rho = t2.sum(-1)
    r = np.sqrt(rho)
    e = np.exp(-math.sqrt(5.0) * r)
    m2 = magn * magn
    g = m2 * (1 + math.sqrt(5.0) * r + 5.0 * rho / 3.0) * e
    dg = -(5.0 / 6.0) * m2 * (1 + math.sqrt(5.0) * r) * e
    return g, dg, t2


def energy_grad_matern52_ard(x, y, lscale, magn, sigma, jitter=1e-6, chunk=128):
    """Returns (log evidence, gradient of the energy wrt [log l_1..l_d, log magn, log sigma], seconds).
    The kernel build and the gradient contraction run over row chunks on a thread
    pool (numpy releases the GIL); the chunks' gradient parts are summed in chunk
    order."""
    t0 = time.perf_counter()
    n, d = x.shape
    w = 1.0 / np.asarray(lscale, float) ** 2
    starts = list(range(0, n, chunk))
    pool = ThreadPoolExecutor(max_workers=_threads())
    K = np.empty((n, n))

    def build(s):
        K[s:s + chunk] = _m52_parts(x[s:s + chunk], x, w, magn)[0]

    list(pool.map(build, starts))
    K[np.diag_indices(n)] += jitter
    s2 = sigma * sigma
    B = K / s2
    B[np.diag_indices(n)] += 1.0
    Lb, info = lapack.dpotrf(B, lower=1, overwrite_a=1, clean=1)
    if info != 0:
        raise np.linalg.LinAlgError(f"dpotrf info={info}")
    sv = 1.0 / sigma
    u = sla.solve_triangular(Lb, sv * y, lower=True, check_finite=False)
    alpha = sv * sla.solve_triangular(Lb, u, lower=True, trans="T", check_finite=False)
    f = K @ alpha
    ll = float(np.sum(-0.5 * math.log(2 * math.pi) - math.log(sigma) - 0.5 * ((y - f) / sigma) ** 2))
    logz = ll - 0.5 * float(f @ alpha) - float(np.sum(np.log(np.diag(Lb))))
    del K
    Binv, info = lapack.dpotri(Lb, lower=1, overwrite_c=1)
    if info != 0:
        raise np.linalg.LinAlgError(f"dpotri info={info}")
    # Binv holds the lower triangle; symmetrise
    Binv = np.tril(Binv) + np.tril(Binv, -1).T
    def grad_part(s):
        gk, dg, t2 = _m52_parts(x[s:s + chunk], x, w, magn)
        coef = alpha[s:s + chunk, None] * alpha[None, :] - Binv[s:s + chunk] / s2
        cd = -2.0 * coef * dg
        gp = np.empty(d + 1)
        gp[:d] = np.einsum("ij,ijk->k", cd, t2)
        gp[d] = float(np.sum(2.0 * coef * gk))
        return gp

    g = np.zeros(d + 1)
    for gp in pool.map(grad_part, starts):  # results in chunk order
        g += gp
    pool.shutdown()
    grad = np.empty(d + 2)
    grad[:d + 1] = -0.5 * g
    grad[d + 1] = -s2 * float(alpha @ alpha) + float(np.trace(Binv))
    return logz, grad, time.perf_counter() - t0
