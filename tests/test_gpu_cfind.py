"""cfind in gp_pred and gp_draw (R/gp_pred.R:88-122,206-207, R/gp_draw.R:3-35):
prediction with a subset of the covariance functions on the device
(gpl_pred_terms, gpl_draw_terms) against the oracle with the same subset,
relative 1e-8, for a model with a sum, a product (two device terms) and a
constant; the prior (unfitted) path; and argument checks."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _model(G, lik="gaussian"):
    r = np.random.default_rng(17)
    n = 160
    x = r.normal(size=(n, 3))
    f = np.sin(2 * x[:, 0]) + 0.5 * x[:, 1] * np.cos(x[:, 2])
    cfs = [G.cf_sexp(vars=[0], lscale=0.7, magn=1.0),
           G.cf_matern32(vars=[1, 2], lscale=[1.1, 0.9], magn=0.8) * G.cf_lin(vars=[1], prior_magn=G.prior_fixed()),
           G.cf_const(magn=0.4)]
    if lik == "gaussian":
        lk, y, kw = G.lik_gaussian(sigma=0.2), f + 0.2 * r.normal(size=n), dict(maxiter=50, tol=1e-4)
    else:
        lk = G.lik_bernoulli("logit")
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-2 * f))).astype(float)
        kw = dict(maxiter=50, tol=1e-8)
    gp = G.gp_init(cfs, lk, approx=G.approx_laplace(maxiter=50, tol=1e-8))
    gf = G.gp_fit(gp, x, y)
    fo = O.gp_fit([c.spec() for c in gf.cfs], lk.spec(), x, y, **kw)
    return gp, gf, fo, x, y


@pytest.mark.parametrize("lik", ["gaussian", "logit"])
@pytest.mark.parametrize("cfind", [[0], [1], [2, 0], [0, 1, 2]])
def test_pred_cfind_matches_oracle(gpl, lik, cfind):
    G = gpl
    _, gf, fo, _, _ = _model(G, lik)
    xt = np.random.default_rng(3).normal(size=(90, 3))
    pr = G.gp_pred(gf, xt, var=True, cfind=cfind)
    mo, vo = O.gp_pred(fo, xt, var=True, cfind=cfind)
    assert _rel(pr["mean"], mo) < 1e-8
    assert _rel(pr["var"], vo) < 1e-8
    d = G.gp_draw(gf, xt, draws=2, transform=False, cfind=cfind, seed=4)
    z = np.random.default_rng(4).standard_normal((90, 2))
    assert _rel(d, O.gp_draw(fo, xt, z, cfind=cfind)) < 1e-8


def test_pred_cfind_additive_and_prior(gpl):
    G = gpl
    gp, gf, _, _, _ = _model(G)
    xt = np.random.default_rng(5).normal(size=(40, 3))
    full = G.gp_pred(gf, xt)["mean"]
    parts = sum(G.gp_pred(gf, xt, cfind=k)["mean"] for k in range(3))
    assert _rel(parts, full) < 1e-12
    # prior: variance of the selected functions only (R/gp_pred.R:142-156)
    v = G.gp_pred(gp, xt, var=True, cfind=[1])["var"]
    ref = O.eval_cf([c.spec() for c in gp.cfs], xt, xt, diag_only=True, cfind=[1])
    assert _rel(v, ref) < 1e-12


def test_pred_terms_rejects_bad_index(gpl):
    G = gpl
    from gplite_amd import _lib as L
    import ctypes as C
    from gplite_amd import api
    _, gf, _, _, _ = _model(G)
    xt = np.zeros((4, 3))
    _, _, xtp, _ = api._program(gf.cfs, xt)  # the model's prepared input layout (4 columns)
    mean = np.empty(4)
    bad = (C.c_int32 * 1)(7)
    with pytest.raises(L.GplError) as ei:
        L.check(L.lib().gpl_pred_terms(gf.fit.handle, bad, 1, L.dptr(xtp), 4, 0, L.dptr(mean), None))
    assert ei.value.code == L.GPL_ERR_ARG
    with pytest.raises(IndexError):
        G.gp_pred(gf, xt, cfind=[3])
