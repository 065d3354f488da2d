"""The iteration cap, as the reference applies it.

R/laplace.R:293-295 warns and R/ep.R:322-325 sets log_evidence = -Inf whenever
the loop index equals maxiter after the loop, which includes a loop whose last
allowed step converged.  The oracle is run first to find the step count k at
which the loop converges; the device fit with maxiter = k must then report the
cap (Laplace: a warning, the fit kept; EP: log evidence -Inf) exactly as the
oracle does, and with maxiter = k + 1 it must not.
"""

import warnings

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _data(n, seed):
    r = np.random.default_rng(seed)
    x = r.normal(size=(n, 2))
    f = np.sin(2 * x[:, 0]) + 0.5 * x[:, 1]
    y = (r.uniform(size=n) < 1 / (1 + np.exp(-2 * f))).astype(float)
    return x, y


def _oracle_fit(gp, x, y, approx, maxiter):
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        fo = O.gp_fit([c.spec() for c in gp.cfs], gp.lik.spec(), x, y, approx=approx, maxiter=maxiter)
    return fo, [str(m.message) for m in w if "Maximum number of iterations" in str(m.message)]


def _device_fit(G, gp, x, y, prec):
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        gf = G.gp_fit(gp, x, y, precision=prec)
    return gf, [str(m.message) for m in w if "Maximum number of iterations" in str(m.message)]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_laplace_cap_on_the_converging_step(gpl, prec):
    G = gpl
    x, y = _data(150, 11)
    cf = G.cf_sexp(lscale=0.8, magn=1.5)
    free, _ = _oracle_fit(G.gp_init(cf, G.lik_bernoulli("logit")), x, y, "laplace", 30)
    k = free["iters"]
    assert 2 <= k < 30
    for cap, capped in ((k, True), (k + 1, False)):
        gp = G.gp_init(G.cf_sexp(lscale=0.8, magn=1.5), G.lik_bernoulli("logit"),
                       approx=G.approx_laplace(maxiter=cap))
        fo, wo = _oracle_fit(gp, x, y, "laplace", cap)
        gf, wd = _device_fit(G, gp, x, y, prec)
        assert bool(wo) == capped and bool(wd) == capped, (cap, wo, wd)
        assert gf.fit.iters == fo["iters"] == k
        # the fit is kept when the cap only warns
        tol = 1e-8 if prec == "fp64" else 1e-6
        assert gf.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=tol)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_ep_cap_on_the_converging_sweep(gpl, prec):
    G = gpl
    x, y = _data(120, 12)
    free, _ = _oracle_fit(G.gp_init(G.cf_sexp(lscale=0.8, magn=1.5), G.lik_bernoulli("probit"),
                                    approx=G.approx_ep()), x, y, "ep", 100)
    k = free["iters"]
    assert 2 <= k < 100
    for cap, capped in ((k, True), (k + 1, False)):
        gp = G.gp_init(G.cf_sexp(lscale=0.8, magn=1.5), G.lik_bernoulli("probit"),
                       approx=G.approx_ep(maxiter=cap))
        fo, wo = _oracle_fit(gp, x, y, "ep", cap)
        gf, wd = _device_fit(G, gp, x, y, prec)
        assert bool(wo) == capped and bool(wd) == capped, (cap, wo, wd)
        assert gf.fit.iters == fo["iters"] == k
        if capped:
            assert fo["log_evidence"] == -np.inf
            assert gf.fit.log_evidence == -np.inf
        else:
            tol = 1e-8 if prec == "fp64" else 1e-6 * len(y)
            assert np.isfinite(gf.fit.log_evidence)
            assert abs(gf.fit.log_evidence - fo["log_evidence"]) <= tol * max(1.0, abs(fo["log_evidence"]))
