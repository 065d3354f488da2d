"""The beta-binomial likelihood and the non-log-concave branch on the device.

Beta-binomial outliers (y = 0 of 10 among neighbours near 10 of 10) make the log
density locally convex at the mode, so some pseudo-variances are negative and
the Laplace step and the EP sweep split their sites (R/laplace.R:40-117,
R/ep.R:50-138).  The oracle restates the reference's split with its final LU
solve; the device runs block elimination over the same split.  Both are held to
1e-8 relative: log evidence, step count, posterior mode / mean, alpha, EP
posterior variances, and the predictive mean and variance (the reference's
C_lu branch of gp_pred, R/gp_pred.R:220-223).
"""

import warnings

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-8


def rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _data(n, every=None, outl=(10, 50, 90), seed=3, phi=0.2):
    r = np.random.default_rng(seed)
    x = np.sort(r.uniform(-3, 3, n))
    f = 3.0 + 0.5 * np.sin(x)
    mu = 1 / (1 + np.exp(-f))
    y = r.binomial(10, r.beta(mu / phi, (1 - mu) / phi)).astype(float)
    if every:
        y[::every] = 0.0
    else:
        y[list(outl)] = 0.0
    return x[:, None], y, np.full(n, 10.0)


def _oracle(gp, x, y, t, approx):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return O.gp_fit([c.spec() for c in gp.cfs], gp.lik.spec(), x, y, approx=approx, trials=t)


@pytest.mark.parametrize("n,every,link", [(100, None, "logit"), (700, 10, "logit"), (300, 15, "probit")])
@pytest.mark.parametrize("approx", ["laplace", "ep"])
def test_split_fit_and_prediction_match_oracle(gpl, n, every, link, approx):
    G = gpl
    x, y, t = _data(n, every)
    gp = G.gp_init(G.cf_sexp(lscale=1.0, magn=1.0), G.lik_betabinom(link, phi=0.2),
                   approx=G.approx_ep() if approx == "ep" else G.approx_laplace())
    fo = _oracle(gp, x, y, t, approx)
    assert "C_lu" in fo, "the case must reach the split branch"
    nbad = int(np.sum(fo["V"] < 0)) if "V" in fo else int(np.sum(fo["P"] < 0))
    gf = G.gp_fit(gp, x, y, trials=t)
    print(f"{approx} n={n} link={link}: {nbad} split sites, steps {gf.fit.iters} vs {fo['iters']}, "
          f"logZ {gf.fit.log_evidence!r} vs {fo['log_evidence']!r}", flush=True)
    assert gf.fit.iters == fo["iters"]
    assert gf.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=RTOL)
    assert rel(gf.fit.fmean, fo["fmean"]) < RTOL
    assert rel(gf.fit.alpha, fo["alpha"]) < RTOL
    if approx == "ep":
        assert rel(gf.fit.fvar, fo["fvar"]) < RTOL
    xt = np.linspace(-3.5, 3.5, 129)[:, None]
    pr = G.gp_pred(gf, xt, var=True)
    mo, vo = O.gp_pred(fo, xt, var=True)
    assert rel(pr["mean"], mo) < RTOL
    assert rel(pr["var"], vo) < RTOL
    # the prediction at the training inputs (the Laplace LOO's variance)
    pr = G.gp_pred(gf, x, var=True)
    mo, vo = O.gp_pred(fo, x, var=True)
    assert rel(pr["var"], vo) < RTOL


@pytest.mark.parametrize("approx", ["laplace", "ep"])
def test_mixed_precision_split_fit_is_refitted_in_fp64(gpl, approx):
    """A split site makes the mixed form redo the fit in FP64, which splits on the device."""
    G = gpl
    x, y, t = _data(100)
    gp = G.gp_init(G.cf_sexp(lscale=1.0, magn=1.0), G.lik_betabinom(phi=0.2),
                   approx=G.approx_ep() if approx == "ep" else G.approx_laplace())
    gm = G.gp_fit(gp, x, y, trials=t, precision="mixed")
    assert gm.fit.precision_info["refit_fp64"]
    gd = G.gp_fit(gp, x, y, trials=t)
    assert gm.fit.log_evidence == gd.fit.log_evidence
    assert np.array_equal(gm.fit.fmean, gd.fit.fmean)


@pytest.mark.parametrize("approx", ["laplace", "ep"])
def test_betabinom_without_outliers_matches_oracle(gpl, approx):
    """The beta-binomial density, gradient and EP quadrature on the ordinary path."""
    G = gpl
    r = np.random.default_rng(5)
    x = r.uniform(-3, 3, (150, 2))
    f = np.sin(x[:, 0]) + 0.5 * x[:, 1]
    t = r.integers(1, 20, 150).astype(float)
    mu = 1 / (1 + np.exp(-f))
    y = r.binomial(t.astype(int), r.beta(mu / 0.5, (1 - mu) / 0.5)).astype(float)
    gp = G.gp_init(G.cf_matern32(lscale=1.0, magn=1.2), G.lik_betabinom(phi=0.5),
                   approx=G.approx_ep() if approx == "ep" else G.approx_laplace())
    fo = _oracle(gp, x, y, t, approx)
    assert "C_lu" not in fo
    gf = G.gp_fit(gp, x, y, trials=t)
    assert gf.fit.iters == fo["iters"]
    assert gf.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=RTOL)
    assert rel(gf.fit.fmean, fo["fmean"]) < RTOL
    xt = r.uniform(-3, 3, (64, 2))
    pr = G.gp_pred(gf, xt, var=True)
    mo, vo = O.gp_pred(fo, xt, var=True)
    assert rel(pr["mean"], mo) < RTOL and rel(pr["var"], vo) < RTOL


def test_split_fit_has_no_cholesky_factor(gpl):
    """Like the reference's C_lu fit, a split fit has no C_chol and no draws on the device."""
    G = gpl
    from gplite_amd import _lib as L
    x, y, t = _data(100)
    gf = G.gp_fit(G.gp_init(G.cf_sexp(lscale=1.0, magn=1.0), G.lik_betabinom(phi=0.2)), x, y, trials=t)
    with pytest.raises(L.GplError) as ei:
        _ = gf.fit.C_chol
    assert ei.value.code == L.GPL_ERR_UNSUPPORTED
    with pytest.raises(L.GplError):
        G.gp_draw(gf, x[:5], draws=3)
