"""Mixed-precision form (FP32 MFMA storage/updates, FP64 diagonal blocks,
log-determinant and refinement) against the FP64 oracle.

Stated tolerances:
* log evidence: 1e-6 relative to the size of its summands (logz_tol below);
* alpha and the predictive mean: 1e-6 relative (they come from FP64 refinement);
* gradient: 1e-3 relative (it comes from an FP32 B^-1);
* predictive variance: 1e-3 relative (var = k** - q with q close to k**, and q
  comes from the FP32 inverse factor).
"""

import math

import numpy as np
import pytest

import oracle as O
from tests.test_gpu_core import GRAD_CASES

pytestmark = pytest.mark.gpu


def logz_tol(n, sigma, logz, rtol=1e-6):
    """Absolute bound for the mixed form's log evidence.

    The FP32 trailing updates of the Cholesky factor perturb log|B| by roughly
    sqrt(n) * 2^-24 * max_i B_ii in absolute terms, whatever |logZ| is.  logZ is
    a sum whose normalising terms have size n(|log sigma| + log(2 pi)/2), so
    rtol is applied to the larger of |logZ| and that size.  When the terms do
    not cancel this is the same as rtol relative to |logZ|.
    """
    return rtol * max(abs(logz), n * (abs(math.log(sigma)) + 0.5 * math.log(2 * math.pi)))


def rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("n,d,sigma", [(64, 1, 0.1), (333, 2, 0.1), (1500, 8, 0.05), (2048, 8, 0.05)])
def test_mixed_gaussian_parity(gpl, n, d, sigma):
    G = gpl
    r = np.random.default_rng(n + d)
    x = r.uniform(0, 1, (n, d))
    y = np.sin(3 * x.sum(1)) + sigma * r.normal(size=n)
    ls = r.uniform(0.2, 1.0, d)
    gp = G.gp_init(G.cf_matern52(lscale=ls, magn=1.0), G.lik_gaussian(sigma=sigma))
    gm = G.gp_fit(gp, x, y, precision="mixed")
    fo = O.gp_fit([c.spec() for c in gm.cfs], gm.lik.spec(), x, y)
    assert abs(gm.fit.log_evidence - fo["log_evidence"]) <= logz_tol(n, sigma, fo["log_evidence"])
    assert rel(gm.fit.alpha, fo["alpha"]) < 1e-6
    xt = r.uniform(0, 1, (257, d))
    pr = G.gp_pred(gm, xt, var=True)
    mo, vo = O.gp_pred(fo, xt, var=True)
    assert rel(pr["mean"], mo) < 1e-6
    g = G.gp_energy_grad(gm)
    go = O.gp_energy_grad(fo)
    assert rel(g, go) < 1e-3
    pr = G.gp_pred(gm, xt, var=True)
    assert rel(pr["var"], vo) < 1e-3


def test_mixed_matches_plain_device_path(gpl):
    """The two device forms agree on the same inputs (FP64 plain as reference)."""
    G = gpl
    r = np.random.default_rng(5)
    n = 1024
    x = r.normal(size=(n, 3)); y = np.cos(x.sum(1)) + 0.1 * r.normal(size=n)
    gp = G.gp_init(G.cf_sexp(lscale=[0.6, 0.8, 1.0], magn=1.3), G.lik_gaussian(sigma=0.1))
    g64 = G.gp_fit(gp, x, y, want_grad=True)
    g32 = G.gp_fit(gp, x, y, precision="mixed")
    assert abs(g32.fit.log_evidence - g64.fit.log_evidence) <= logz_tol(n, 0.1, g64.fit.log_evidence)
    assert rel(G.gp_energy_grad(g32), G.gp_energy_grad(g64)) < 1e-3


def test_mixed_n8192_workload_meets_1e6(gpl):
    """The n=8192, d=8 Matern-5/2 ARD regression: the mixed form's log evidence and
    predictive mean within 1e-6 relative of the FP64 device path (itself held to
    1e-8 of the oracle on the other tests)."""
    G = gpl
    from gplite_amd import synth
    w = synth.c2()
    gp = G.gp_init(G.cf_matern52(lscale=w["lscale"], magn=w["magn"]), G.lik_gaussian(sigma=w["sigma"]))
    g64 = G.gp_fit(gp, w["x"], w["y"])
    g32 = G.gp_fit(gp, w["x"], w["y"], precision="mixed")
    assert abs(g32.fit.log_evidence / g64.fit.log_evidence - 1) <= 1e-6
    xt = w["xt"][:1024]
    assert rel(G.gp_pred(g32, xt)["mean"], G.gp_pred(g64, xt)["mean"]) <= 1e-6


@pytest.mark.parametrize("lik_name,n", [("logit", 400), ("probit", 700), ("poisson", 1000)])
def test_mixed_ep_parity(gpl, lik_name, n):
    """Parallel EP in the mixed form (FP32 factor and L^-1, FP64 diagonal tiles and
    refinement): the sweep count, posterior mean and predictive mean follow the FP64
    oracle to 1e-6 relative, the posterior variance to 1e-5, the log evidence to
    1e-6 relative to the size of its summands (n)."""
    G = gpl
    r = np.random.default_rng(n + 17)
    x = r.normal(size=(n, 3))
    f = np.sin(2 * x[:, 0]) + 0.5 * x[:, 1]
    if lik_name == "poisson":
        lk = G.lik_poisson()
        y = r.poisson(np.exp(0.5 * f)).astype(float)
    else:
        lk = G.lik_bernoulli(lik_name)
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-f))).astype(float)
    gp = G.gp_init(G.cf_sexp(lscale=[0.8, 0.9, 1.2], magn=1.5), lk, approx=G.approx_ep())
    gm = G.gp_fit(gp, x, y, precision="mixed")
    fo = O.gp_fit([c.spec() for c in gm.cfs], lk.spec(), x, y, approx="ep")
    assert gm.fit.iters == fo["iters"]
    assert abs(gm.fit.log_evidence - fo["log_evidence"]) <= 1e-6 * max(abs(fo["log_evidence"]), n)
    assert rel(gm.fit.fmean, fo["fmean"]) < 1e-6
    assert rel(gm.fit.fvar, fo["fvar"]) < 1e-5
    xt = r.normal(size=(64, 3))
    pr = G.gp_pred(gm, xt, var=True)
    mo, vo = O.gp_pred(fo, xt, var=True)
    assert rel(pr["mean"], mo) < 1e-6
    assert rel(pr["var"], vo) < 1e-3


@pytest.mark.parametrize("lik_name,n", [("logit", 400), ("poisson", 1000)])
def test_mixed_laplace_parity(gpl, lik_name, n):
    """Laplace Newton loop in the mixed form: the iterate sequence (step count),
    posterior mode and predictive mean follow the FP64 oracle to 1e-6 relative,
    the log evidence to 1e-6 relative to the size of its summands (n)."""
    G = gpl
    r = np.random.default_rng(n + 7)
    x = r.normal(size=(n, 3))
    f = np.sin(2 * x[:, 0]) + 0.5 * x[:, 1]
    if lik_name == "poisson":
        lk = G.lik_poisson()
        y = r.poisson(np.exp(0.5 * f)).astype(float)
    else:
        lk = G.lik_bernoulli(lik_name)
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-f))).astype(float)
    gp = G.gp_init(G.cf_sexp(lscale=[0.8, 0.9, 1.2], magn=1.5), lk, approx=G.approx_laplace(maxiter=50, tol=1e-8))
    gm = G.gp_fit(gp, x, y, precision="mixed")
    fo = O.gp_fit([c.spec() for c in gm.cfs], lk.spec(), x, y, maxiter=50, tol=1e-8)
    assert gm.fit.iters == fo["iters"]
    assert abs(gm.fit.log_evidence - fo["log_evidence"]) <= 1e-6 * max(abs(fo["log_evidence"]), n)
    assert rel(gm.fit.fmean, fo["fmean"]) < 1e-6
    xt = r.normal(size=(64, 3))
    pr = G.gp_pred(gm, xt, var=True)
    mo, vo = O.gp_pred(fo, xt, var=True)
    assert rel(pr["mean"], mo) < 1e-6
    assert rel(pr["var"], vo) < 1e-3


@pytest.mark.parametrize("ci", range(len(GRAD_CASES)))
def test_mixed_gradient_composites(gpl, ci):
    """The mixed form's gradient for every covariance structure of the FP64
    gradient tests (sums, products, nested products, periodic, nn, lin), which
    go through the generic gradient kernel on the FP32 B^-1: 1e-3 relative to
    the oracle, the log evidence to 1e-6 of its summands."""
    G = gpl
    r = np.random.default_rng(100 + ci)
    n = 700
    x = r.normal(size=(n, 2))
    y = np.sin(x[:, 0]) * np.cos(0.5 * x[:, 1]) + 0.25 * r.normal(size=n)
    gp = G.gp_init(GRAD_CASES[ci](G), G.lik_gaussian(sigma=0.25, prior_sigma=G.prior_lognormal(-1.0, 0.5)))
    gm = G.gp_fit(gp, x, y, precision="mixed")
    fo = O.gp_fit([c.spec() for c in gm.cfs], gm.lik.spec(), x, y)
    assert abs(gm.fit.log_evidence - fo["log_evidence"]) <= logz_tol(n, 0.25, fo["log_evidence"])
    assert rel(G.gp_energy_grad(gm), O.gp_energy_grad(fo)) < 1e-3
