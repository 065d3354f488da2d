"""The reference's prediction tests (tests/testthat/test_pred.R:85-141 and
tests/testthat/test_sparse.R:98-172) on the device, for method_full and
method_fitc(num_inducing = 10), every covariance function alone and the
Gaussian, Bernoulli, binomial and Poisson likelihoods (lik_betabinom is not
provided):

* prediction and draws work on the prior (method_full; for method_fitc only
  the mean, the inducing inputs being unset), on a fit, and after a refit, and
  raise after set_param until the model is refitted;
* the analytic predictive mean and sd match those of 2e5 joint posterior draws
  (transform = FALSE): the sd in R's all.equal sense, mean |a - b| / mean |b|
  <= 1e-2; the mean to 5 Monte Carlo standard errors (sd / sqrt(draws)) per
  point, because R's relative form divides by mean |b|, which for posterior
  means near 0 (Bernoulli with magn = 0.1) is of the order of the Monte Carlo
  error itself.

Data follow the reference's construction (n = 30, x ~ U(-3, 3), 5 test points,
f = 2x - 2, trials ~ U{1..10}) from numpy's generator."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(G):
    r = np.random.default_rng(1234)
    n, nt = 30, 5
    x = r.uniform(size=n) * 6 - 3
    xt = r.uniform(size=nt) * 6 - 3
    f = 2 * x - 2
    trials = r.integers(1, 11, size=n).astype(float)
    cfs = [lambda: G.cf_const(magn=0.1), lambda: G.cf_lin(magn=0.1), lambda: G.cf_sexp(magn=0.1),
           lambda: G.cf_matern32(magn=0.1), lambda: G.cf_matern52(magn=0.1), lambda: G.cf_nn(magn=0.1),
           lambda: G.cf_periodic()]
    liks = [lambda: G.lik_gaussian(), lambda: G.lik_bernoulli("logit"), lambda: G.lik_binomial("logit"),
            lambda: G.lik_poisson()]
    return r, x, xt, f, trials, cfs, liks


def _target(r, lik, f, trials):
    """generate_target (R/lik.R:504-537)."""
    if lik.kind == "gaussian":
        return f + lik.sigma * r.standard_normal(len(f))
    mu = lik.response(f)
    if lik.kind == "poisson":
        return r.poisson(mu).astype(float)
    t = np.ones(len(f)) if lik.kind == "bernoulli" else trials
    return r.binomial(t.astype(int), mu).astype(float)


def _all_equal(a, b, tol=1e-2):
    return float(np.mean(np.abs(a - b)) / np.mean(np.abs(b))) <= tol


def _models(G, method):
    r, x, xt, f, trials, cfs, liks = _setup(G)
    for li, lk in enumerate(liks):
        for ci, cf in enumerate(cfs):
            if method == "fitc" and ci in (3, 4):  # the reference's sparse test leaves out the Matern kernels
                continue
            meth = G.method_fitc(num_inducing=10) if method == "fitc" else G.method_full()
            gp = G.gp_init(cf(), lk(), method=meth)
            y = _target(r, gp.lik, f, trials)
            yield f"{lk().kind}-{ci}", gp, x, xt, y, trials


@pytest.mark.parametrize("method", ["full", "fitc"])
def test_refit_guard_with_draws(gpl, method):
    G = gpl
    for name, gp0, x, _, y, trials in _models(G, method):
        kw = {"trials": trials} if gp0.lik.kind == "binomial" else {}
        if method == "full":
            G.gp_draw(gp0, x, draws=1)
            G.gp_pred(gp0, x, var=True)
        else:
            with pytest.raises(ValueError):
                G.gp_draw(gp0, x, draws=1)
            with pytest.raises(ValueError):
                G.gp_pred(gp0, x, var=True)
        G.gp_pred(gp0, x, var=False)
        gp = G.gp_fit(gp0, x, y, **kw)
        G.gp_draw(gp, x, draws=1)
        G.gp_pred(gp, x, var=True)
        G.gp_pred(gp, x, var=False)
        p = np.array(list(G.get_param(gp).values()))
        p[0] = -2.0 if method == "fitc" else 0.0
        gp = G.set_param(gp, p)
        for call in (lambda: G.gp_draw(gp, x, draws=1), lambda: G.gp_pred(gp, x, var=True),
                     lambda: G.gp_pred(gp, x, var=False)):
            with pytest.raises(RuntimeError):
                call()
        gp1 = G.gp_fit(gp, x, y, **kw)
        G.gp_draw(gp1, x, draws=1)
        G.gp_pred(gp1, x, var=True)


@pytest.mark.parametrize("method", ["full", "fitc"])
def test_analytic_matches_sampling(gpl, method):
    G = gpl
    for name, gp0, x, xt, y, trials in _models(G, method):
        kw = {"trials": trials} if gp0.lik.kind == "binomial" else {}
        gp = G.gp_fit(gp0, x, y, **kw)
        pred = G.gp_pred(gp, xt, var=True)
        nd = 200000
        draws = G.gp_draw(gp, xt, draws=nd, transform=False, seed=4321)
        sd = np.sqrt(pred["var"])
        assert np.all(np.abs(draws.mean(axis=1) - pred["mean"]) <= 5 * sd / np.sqrt(nd)), name
        assert _all_equal(draws.std(axis=1, ddof=1), sd), name
