"""gp_draw (R/gp_draw.R:101-124, method_full) on the device fit: joint posterior
draws mean + chol(cov + jitter I) z against the oracle's restatement on the same
standard normal matrix z, for the Gaussian and Laplace (Bernoulli-logit) fits,
ragged test sizes (1, 65, 130 points) and several draw counts.

* FP64 path: draws relative 1e-8 to the oracle (max |error| / max |draw|).
* z = I gives the factor itself: L L^T must reproduce the oracle's covariance
  (FP64 1e-10; mixed precision 2e-5 absolute with jitter 1e-4: the mixed fit
  keeps L^-1 in FP32, so its covariance carries an error near 1e-6 that a
  jitter of 1e-6 does not cover).
* marginal = TRUE uses the predictive variances (R/gp_draw.R:104-107).
* A covariance that is not positive definite is reported as NOT_PD."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _data(case, n=150, seed=31):
    r = np.random.default_rng(seed)
    x = r.normal(size=(n, 2))
    f = np.sin(x.sum(1))
    if case == "gaussian":
        y = f + 0.3 * r.normal(size=n)
    else:
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-2 * f))).astype(float)
    return x, y


def _fit(G, case, x, y, precision=None):
    if case == "gaussian":
        lk, approx = G.lik_gaussian(sigma=0.3), G.approx_laplace()
    else:
        lk, approx = G.lik_bernoulli("logit"), G.approx_laplace(maxiter=50, tol=1e-8)
    gp = G.gp_init(G.cf_sexp(lscale=[0.9, 1.1], magn=1.2), lk, approx=approx)
    gf = G.gp_fit(gp, x, y, precision=precision)
    kw = dict(maxiter=50, tol=1e-8 if case != "gaussian" else 1e-4)
    fo = O.gp_fit([c.spec() for c in gf.cfs], lk.spec(), x, y, **kw)
    return gf, fo


@pytest.mark.parametrize("case", ["gaussian", "logit_laplace"])
@pytest.mark.parametrize("m,draws", [(1, 1), (65, 3), (130, 7)])
def test_draw_matches_oracle(gpl, case, m, draws):
    G = gpl
    x, y = _data(case)
    gf, fo = _fit(G, case, x, y)
    xt = np.random.default_rng(m).normal(size=(m, 2))
    d = G.gp_draw(gf, xt, draws=draws, transform=False, seed=5)
    z = np.random.default_rng(5).standard_normal((m, draws))
    ref = O.gp_draw(fo, xt, z)
    assert d.shape == (m, draws)
    assert np.max(np.abs(d - ref)) <= 1e-8 * np.max(np.abs(ref))
    # transform = TRUE: the response of the same draws
    dt = G.gp_draw(gf, xt, draws=draws, seed=5)
    # the same 1e-8 of the draws' scale as above, carried through the response
    # (an element near zero has no relative accuracy of its own)
    np.testing.assert_allclose(dt, O.response(fo["lik"], ref), rtol=1e-7, atol=1e-8 * np.max(np.abs(ref)))


@pytest.mark.parametrize("precision,tol,jitter", [("fp64", 1e-10, 1e-6), ("mixed", 2e-5, 1e-4)])
def test_draw_factor_reproduces_covariance(gpl, precision, tol, jitter):
    G = gpl
    x, y = _data("gaussian", n=300)
    gf, fo = _fit(G, "gaussian", x, y, precision=precision)
    m = 90
    xt = np.random.default_rng(4).normal(size=(m, 2))
    mean, cov = O.gp_pred_cov(fo, xt, jitter=jitter)
    # z = I through the public entry point: the draws minus the mean are L
    from gplite_amd import _lib as L
    from gplite_amd import api
    _, _, xtp, _ = api._program(gf.cfs, xt)
    z = np.asfortranarray(np.eye(m))
    out = np.empty((m, m), order="F")
    L.check(L.lib().gpl_draw(gf.fit.handle, L.dptr(xtp), m, jitter, z.ctypes.data_as(L._dp), m,
                             out.ctypes.data_as(L._dp)))
    Lf = out - mean[:, None]
    assert np.max(np.abs(np.triu(Lf, 1))) <= tol
    assert np.max(np.abs(Lf @ Lf.T - cov)) <= tol


def test_draw_marginal(gpl):
    G = gpl
    x, y = _data("gaussian")
    gf, fo = _fit(G, "gaussian", x, y)
    xt = np.random.default_rng(2).normal(size=(40, 2))
    d = G.gp_draw(gf, xt, draws=4, marginal=True, seed=9)
    mean, v = O.gp_pred(fo, xt, var=True)
    z = np.random.default_rng(9).standard_normal((40, 4))
    ref = mean[:, None] + np.sqrt(v)[:, None] * z
    assert np.max(np.abs(d - ref)) <= 1e-8 * np.max(np.abs(ref))


def test_draw_not_pd_reported(gpl):
    G = gpl
    from gplite_amd import _lib as L
    x, y = _data("gaussian")
    gf, _ = _fit(G, "gaussian", x, y)
    xt = np.random.default_rng(3).normal(size=(10, 2))
    with pytest.raises(L.GplError) as ei:
        G.gp_draw(gf, xt, draws=2, jitter=-10.0, seed=1)
    assert ei.value.code == L.GPL_ERR_NOT_PD


def test_prior_draw_matches_oracle(gpl):
    """gp_draw on an unfitted model (gp_draw_prior, R/gp_draw.R:56-67): chol(K_tt + jitter I) z."""
    G = gpl
    gp = G.gp_init([G.cf_sexp(lscale=0.7, magn=1.3), G.cf_matern32(lscale=1.1, magn=0.4)], G.lik_gaussian())
    xt = np.random.default_rng(8).normal(size=(70, 2))
    d = G.gp_draw(gp, xt, draws=3, transform=False, seed=6)
    z = np.random.default_rng(6).standard_normal((70, 3))
    ref = O.prior_draw([c.spec() for c in gp.cfs], xt, z)
    assert np.max(np.abs(d - ref)) <= 1e-8 * np.max(np.abs(ref))
