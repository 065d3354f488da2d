"""gp_loo (R/gp_loo.R:73-134) on the device fit: LOO predictive log densities
against the oracle's restatement, relative 1e-8 (FP64 path), for the Gaussian,
Laplace (Bernoulli-logit) and EP (Poisson) approximations, with quadrature and
with Monte Carlo draws (same seed on both sides)."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("case", ["gaussian", "logit_laplace", "poisson_ep"])
def test_loo_matches_oracle(gpl, case):
    G = gpl
    r = np.random.default_rng(21)
    n = 150
    x = r.normal(size=(n, 2))
    f = np.sin(x.sum(1))
    if case == "gaussian":
        lk = G.lik_gaussian(sigma=0.3)
        y = f + 0.3 * r.normal(size=n)
        approx = G.approx_laplace()
    elif case == "logit_laplace":
        lk = G.lik_bernoulli("logit")
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-2 * f))).astype(float)
        approx = G.approx_laplace(maxiter=50, tol=1e-8)
    else:
        lk = G.lik_poisson()
        y = r.poisson(np.exp(0.5 * f)).astype(float)
        approx = G.approx_ep()
    gp = G.gp_init(G.cf_sexp(lscale=[0.9, 1.1], magn=1.2), lk, approx=approx)
    gf = G.gp_fit(gp, x, y)
    kw = dict(approx="ep") if case == "poisson_ep" else dict(maxiter=50, tol=1e-8 if case != "gaussian" else 1e-4)
    fo = O.gp_fit([c.spec() for c in gf.cfs], lk.spec(), x, y, **kw)
    lo = O.gp_loo(fo, lk.spec(), x, y, quad_order=11)
    ld = G.gp_loo(gf, x, y, quad_order=11)
    assert _rel(ld["loos"], lo["loos"]) < 1e-8
    assert ld["loo"] == pytest.approx(lo["loo"], rel=1e-8)
    assert ld["sd"] == pytest.approx(lo["sd"], rel=1e-6)
    # Monte Carlo variant: the same normal draws on both sides
    mc_d = G.gp_loo(gf, x, y, quadrature=False, draws=500, seed=3)
    mc_o = O.gp_loo(fo, lk.spec(), x, y, quadrature=False, draws=500, seed=3)
    assert _rel(mc_d["loos"], mc_o["loos"]) < 1e-8
