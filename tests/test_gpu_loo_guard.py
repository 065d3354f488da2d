"""The reference's LOO test (tests/testthat/test_loo.R:114-140) on the device:
for method_full and method_fitc(num_inducing = 10), approx_laplace and
approx_ep(damping = 0.8) (EP not for the Gaussian likelihood), the Gaussian,
Bernoulli, binomial and Poisson likelihoods (lik_betabinom is not provided)
and each covariance function alone, gp_loo raises on an unfitted model,
returns finite values on a fit (quad_order = 7), raises after set_param and
works again after a refit.  Data follow the reference's construction (n = 30,
x ~ U(-3, 3), f = 2x - 4, trials ~ U{1..10}) from numpy's generator."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _target(r, lik, f, trials):
    if lik.kind == "gaussian":
        return f + lik.sigma * r.standard_normal(len(f))
    mu = lik.response(f)
    if lik.kind == "poisson":
        return r.poisson(mu).astype(float)
    t = np.ones(len(f)) if lik.kind == "bernoulli" else trials
    return r.binomial(t.astype(int), mu).astype(float)


@pytest.mark.parametrize("method", ["full", "fitc"])
@pytest.mark.parametrize("approx", ["laplace", "ep"])
def test_loo_guards(gpl, method, approx):
    G = gpl
    r = np.random.default_rng(1234)
    n = 30
    x = r.uniform(size=n) * 6 - 3
    f = 2 * x - 4
    trials = r.integers(1, 11, size=n).astype(float)
    cfs = [G.cf_const, G.cf_lin, G.cf_sexp, G.cf_nn, G.cf_periodic]
    liks = [lambda: G.lik_gaussian(), lambda: G.lik_bernoulli("logit"), lambda: G.lik_binomial("logit"),
            lambda: G.lik_poisson()]
    for lkf in liks:
        lk = lkf()
        if approx == "ep" and lk.kind == "gaussian":
            continue
        for cf in cfs:
            meth = G.method_fitc(num_inducing=10) if method == "fitc" else G.method_full()
            ap = G.approx_ep(damping=0.8) if approx == "ep" else G.approx_laplace()
            gp0 = G.gp_init(cf(), lkf(), method=meth, approx=ap)
            y = _target(r, lk, f, trials)
            kw = {"trials": trials} if lk.kind == "binomial" else {}
            with pytest.raises(ValueError):
                G.gp_loo(gp0, x, y, **kw)
            gp = G.gp_fit(gp0, x, y, **kw)
            loo = G.gp_loo(gp, x, y, quad_order=7, **kw)
            assert np.isfinite(loo["loo"]) and np.all(np.isfinite(loo["loos"])), (method, approx, lk.kind, cf)
            p = np.array(list(G.get_param(gp).values()))
            p[0] = 0.0
            gp = G.set_param(gp, p)
            with pytest.raises(RuntimeError):
                G.gp_loo(gp, x, y, **kw)
            gp1 = G.gp_fit(gp, x, y, **kw)
            G.gp_loo(gp1, x, y, quad_order=7, **kw)
