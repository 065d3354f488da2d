"""The reference's normalization tests (tests/testthat/test_normalization.R:1-139)
on the device: a model whose covariance functions have normalize = TRUE, fitted
on raw inputs, equals the same model without normalization fitted on inputs
standardised beforehand (mean 0, sample sd 1, the scaling of R's scale()).
Models: every one of cf_lin, cf_sexp, cf_matern32, cf_matern52, cf_nn alone,
every pair as a sum, and every triple as a product; ten perturbations of the
hyperparameters each (normal, sd 0.3 on the log scale), Gaussian likelihood.

The reference asks for 1e-3 on the log evidence and 1e-5 on the predictive
mean and variance (R's all.equal).  The two sides differ only by rounding of the
input standardisation, so the bar here is the device path's own: 1e-8 relative
(the variance relative to max(1, max |var|))."""

import copy
import itertools
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _models(G):
    plain = [G.cf_lin(), G.cf_sexp(), G.cf_matern32(), G.cf_matern52(), G.cf_nn()]
    norm = [G.cf_lin(normalize=True), G.cf_sexp(normalize=True), G.cf_matern32(normalize=True),
            G.cf_matern52(normalize=True), G.cf_nn(normalize=True)]
    cp = copy.deepcopy
    out = [([cp(plain[i])], [cp(norm[i])]) for i in range(5)]
    out += [([cp(plain[i]), cp(plain[j])], [cp(norm[i]), cp(norm[j])]) for i, j in itertools.combinations(range(5), 2)]
    with warnings.catch_warnings():  # products warn about several magnitudes (R's SWO)
        warnings.simplefilter("ignore")
        out += [([cp(plain[i]) * cp(plain[j]) * cp(plain[k])], [cp(norm[i]) * cp(norm[j]) * cp(norm[k])])
                for i, j, k in itertools.combinations(range(5), 3)]
    return out


def test_normalize_matches_prescaled_inputs(gpl):
    G = gpl
    r = np.random.default_rng(1234)
    n, nt = 20, 30
    x = 5 * (r.standard_normal(n) - 1)
    xt = 5 * (r.standard_normal(nt) - 1)
    mu, sd = x.mean(), x.std(ddof=1)
    x_norm, xt_norm = (x - mu) / sd, (xt - mu) / sd
    y = 0.1 * x ** 2 + 0.2 * r.standard_normal(n)
    models = _models(G)
    assert len(models) == 25
    for mi, (cf_plain, cf_norm) in enumerate(models):
        gp = G.gp_init(cf_plain, G.lik_gaussian())
        gpn = G.gp_init(cf_norm, G.lik_gaussian())
        p0 = np.array(list(G.get_param(gp).values()))
        assert p0.size == len(G.get_param(gpn))
        for j in range(10):
            p = p0 + 0.3 * r.standard_normal(p0.size)
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                g1 = G.gp_fit(G.set_param(gp, p), x_norm, y)
                g2 = G.gp_fit(G.set_param(gpn, p), x, y)
            case = (mi, j)
            assert g1.fit.log_evidence == pytest.approx(g2.fit.log_evidence, rel=1e-8), case
            p1 = G.gp_pred(g1, xt_norm, var=True, jitter=1e-6)
            p2 = G.gp_pred(g2, xt, var=True, jitter=1e-6)
            assert _rel(p1["mean"], p2["mean"]) < 1e-8, case
            assert np.max(np.abs(p1["var"] - p2["var"])) <= 1e-8 * max(1.0, np.max(np.abs(p2["var"]))), case
