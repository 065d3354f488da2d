"""gpl_fit_batch: many Gaussian fits in one batched mixed-precision pass, each left
as a resident model.  Every model against the oracle at the mixed form's
tolerances (log evidence 1e-6 of its summands, alpha, posterior and predictive
mean 1e-6, predictive variance and gradient 1e-3); a problem's result does not
depend on the other problems of the call (bitwise); a problem whose FP32 pass
fails is fitted again alone and ends in FP64."""

import numpy as np
import pytest

import oracle as O
from tests.test_gpu_mixed import logz_tol

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _problems(G, sizes, seed=21, sigma=None):
    r = np.random.default_rng(seed)
    gps, xs, ys = [], [], []
    for i, n in enumerate(sizes):
        x = r.uniform(0, 1, (n, 3))
        s = 0.1 if sigma is None else sigma[i]
        ys.append(np.sin(3 * x.sum(1)) + s * r.standard_normal(n))
        xs.append(x)
        gps.append(G.gp_init(G.cf_matern52(lscale=list(r.uniform(0.3, 0.8, 3)), magn=1.0), G.lik_gaussian(sigma=s)))
    return gps, xs, ys


def test_fit_batch_models_match_oracle(gpl):
    G = gpl
    from gplite_amd.batch import gp_fit_batch
    gps, xs, ys = _problems(G, [50, 200, 1500, 64, 777, 1024, 130, 900])
    fits = gp_fit_batch(gps, xs, ys)
    xt = np.random.default_rng(3).uniform(0, 1, (300, 3))
    for g, x, y in zip(fits, xs, ys):
        assert g.fit.precision_info["precision"] == "mixed"
        fo = O.gp_fit([c.spec() for c in g.cfs], g.lik.spec(), x, y)
        assert abs(g.fit.log_evidence - fo["log_evidence"]) <= logz_tol(len(y), 0.1, fo["log_evidence"])
        assert rel(g.fit.alpha, fo["alpha"]) < 1e-6
        assert rel(g.fit.fmean, fo["fmean"]) < 1e-6
        pr = G.gp_pred(g, xt, var=True)
        mo, vo = O.gp_pred(fo, xt, var=True)
        assert rel(pr["mean"], mo) < 1e-6 and rel(pr["var"], vo) < 1e-3
        assert rel(G.gp_energy_grad(g), O.gp_energy_grad(fo)) < 1e-3


def test_fit_batch_result_does_not_depend_on_the_other_problems(gpl):
    G = gpl
    from gplite_amd.batch import gp_fit_batch
    gps, xs, ys = _problems(G, [300, 310, 1100, 90], seed=22)
    together = gp_fit_batch(gps, xs, ys)
    xt = np.random.default_rng(4).uniform(0, 1, (100, 3))
    for i in range(len(gps)):
        alone = gp_fit_batch([gps[i]], [xs[i]], [ys[i]])[0]
        assert alone.fit.log_evidence == together[i].fit.log_evidence
        assert np.array_equal(alone.fit.alpha, together[i].fit.alpha)
        a, b = G.gp_pred(alone, xt, var=True), G.gp_pred(together[i], xt, var=True)
        assert np.array_equal(a["mean"], b["mean"]) and np.array_equal(a["var"], b["var"])


def test_fit_batch_refits_a_failed_problem_alone(gpl):
    """sigma = 1e-3: the batched FP32 pass cannot carry it; it is fitted again alone
    and, its refinement not contracting either, ends in FP64."""
    G = gpl
    from gplite_amd.batch import gp_fit_batch
    r = np.random.default_rng(23)
    gps, xs, ys = [], [], []
    for n, s in ((200, 0.1), (300, 1e-3), (250, 0.2)):  # cond(B) ~ 9e7 for the second
        x = r.uniform(0, 1, (n, 2))
        xs.append(x)
        ys.append(np.sin(3 * x.sum(1)) + s * r.standard_normal(n))
        gps.append(G.gp_init(G.cf_sexp(lscale=0.5, magn=1.0), G.lik_gaussian(sigma=s)))
    fits = gp_fit_batch(gps, xs, ys)
    assert fits[1].fit.precision_info["refit_fp64"]
    for g, x, y in zip(fits, xs, ys):
        fo = O.gp_fit([c.spec() for c in g.cfs], g.lik.spec(), x, y)
        assert g.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=1e-6)
    assert not fits[0].fit.precision_info["refit_fp64"] and not fits[2].fit.precision_info["refit_fp64"]
