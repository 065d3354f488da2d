"""The mixed-precision form's guard: iterative refinement with the FP32 inverse
factor contracts only while cond(B) * 2^-24 is well below 1.  When a solve's
last relative correction stays above 1e-9 (or grows), the single-problem fit is
redone in FP64 and says so (precision_info); the batched form, whose sweep count
is fixed, marks such a problem GPL_ERR_NOCONV with energy +Inf and keeps the
others.

Tolerances: a refitted result is the FP64 device path's result (the same code,
1e-12); against the oracle the bar is 1e-7 on the log evidence and 1e-5 on alpha
and the predictive mean, because at cond(B) ~ 1e8 two FP64 solvers already
differ by about cond(B) * 2^-53 ~ 1e-8.
"""

import warnings

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _gauss(n, sigma, seed=5):
    r = np.random.default_rng(seed)
    x = r.uniform(0, 1, (n, 2))
    y = np.sin(3 * x.sum(1)) + sigma * r.normal(size=n)
    return x, y


def test_ill_conditioned_gaussian_is_refitted_in_fp64(gpl):
    """sigma = 1e-3, n = 300: cond(B) ~ 9e7, so cond(B) 2^-24 ~ 5 and FP32 refinement diverges."""
    G = gpl
    x, y = _gauss(300, 1e-3)
    gp = G.gp_init(G.cf_sexp(lscale=0.5, magn=1.0), G.lik_gaussian(sigma=1e-3))
    gm = G.gp_fit(gp, x, y, precision="mixed")
    info = gm.fit.precision_info
    assert info["refit_fp64"] and info["precision"] == "fp64", info
    gd = G.gp_fit(gp, x, y, precision="fp64")
    assert gm.fit.log_evidence == pytest.approx(gd.fit.log_evidence, rel=1e-12)
    assert rel(gm.fit.alpha, gd.fit.alpha) < 1e-12
    fo = O.gp_fit([c.spec() for c in gm.cfs], gm.lik.spec(), x, y)
    assert gm.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=1e-7)
    assert rel(gm.fit.alpha, fo["alpha"]) < 1e-5
    xt = np.random.default_rng(1).uniform(0, 1, (50, 2))
    assert rel(G.gp_pred(gm, xt)["mean"], O.gp_pred(fo, xt)[0]) < 1e-5
    # the gradient works on the refitted (FP64) model
    g = G.gp_energy_grad(gm)
    assert rel(g, G.gp_energy_grad(gd)) < 1e-10


def test_well_conditioned_gaussian_stays_mixed(gpl):
    G = gpl
    x, y = _gauss(300, 0.1)
    gp = G.gp_init(G.cf_sexp(lscale=0.5, magn=1.0), G.lik_gaussian(sigma=0.1))
    gm = G.gp_fit(gp, x, y, precision="mixed")
    info = gm.fit.precision_info
    assert not info["refit_fp64"] and info["precision"] == "mixed", info
    assert info["refine_rel"] <= 1e-9
    fo = O.gp_fit([c.spec() for c in gm.cfs], gm.lik.spec(), x, y)
    assert rel(gm.fit.alpha, fo["alpha"]) < 1e-6


def test_ill_conditioned_laplace_is_refitted_in_fp64(gpl):
    """Poisson counts near e^9.9 (offset 9.9): W ~ 2e4 and cond(B) ~ 7e7 at the mode."""
    G = gpl
    n = 300
    r = np.random.default_rng(7)
    x = r.uniform(0, 1, (n, 2))
    K = O.cf_sexp_c(x, x, 0.5, 2.0) + 1e-6 * np.eye(n)
    f = np.linalg.cholesky(K) @ r.normal(size=n)
    off = np.full(n, 9.9)
    y = r.poisson(np.exp(off + f)).astype(float)
    gp = G.gp_init(G.cf_sexp(lscale=0.5, magn=2.0), G.lik_poisson())
    gm = G.gp_fit(gp, x, y, offset=off, precision="mixed")
    assert gm.fit.precision_info["refit_fp64"], gm.fit.precision_info
    gd = G.gp_fit(gp, x, y, offset=off, precision="fp64")
    assert gm.fit.iters == gd.fit.iters
    assert gm.fit.log_evidence == pytest.approx(gd.fit.log_evidence, rel=1e-12)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        fo = O.gp_fit([c.spec() for c in gm.cfs], gm.lik.spec(), x, y, offset=off)
    assert gm.fit.iters == fo["iters"]
    assert gm.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=1e-7)
    assert rel(gm.fit.fmean, fo["fmean"]) < 1e-5


@pytest.mark.parametrize("kind", ["gaussian", "laplace", "ep"])
def test_unconverged_refinement_falls_back_for_every_fit(gpl, kind):
    """One refinement sweep can never reach 1e-9 (its correction is the whole
    solution), so every mixed fit takes the FP64 path and equals it."""
    G = gpl
    r = np.random.default_rng(3)
    x = r.normal(size=(200, 2))
    if kind == "gaussian":
        y = np.sin(x.sum(1)) + 0.1 * r.normal(size=200)
        gp = G.gp_init(G.cf_matern32(lscale=0.8), G.lik_gaussian(sigma=0.1))
    else:
        y = (r.uniform(size=200) < 1 / (1 + np.exp(-np.sin(2 * x[:, 0])))).astype(float)
        gp = G.gp_init(G.cf_sexp(lscale=0.8, magn=1.5), G.lik_bernoulli("probit" if kind == "ep" else "logit"),
                       approx=G.approx_ep() if kind == "ep" else G.approx_laplace())
    gm = G.gp_fit(gp, x, y, precision="mixed", refine_iters=1)
    assert gm.fit.precision_info["refit_fp64"]
    gd = G.gp_fit(gp, x, y, precision="fp64")
    assert gm.fit.iters == gd.fit.iters
    assert gm.fit.log_evidence == pytest.approx(gd.fit.log_evidence, rel=1e-13)
    assert rel(gm.fit.fmean, gd.fit.fmean) < 1e-13
    # with the default sweep count the same fits stay in the mixed form
    gm2 = G.gp_fit(gp, x, y, precision="mixed")
    assert not gm2.fit.precision_info["refit_fp64"]


def test_batched_problem_without_convergence_is_flagged(gpl):
    """A batch with one problem at sigma = 1e-3 among well-posed ones: that problem
    is reported (NOCONV, or NOT_PD if its FP32 factor fails) with energy +Inf; the
    others match the oracle."""
    G = gpl
    from gplite_amd import _lib as L
    from gplite_amd.batch import GPBatch
    r = np.random.default_rng(9)
    gps, xs, ys = [], [], []
    sig = [0.1, 1e-3, 0.2, 0.1]
    for i, s in enumerate(sig):
        n = [150, 300, 90, 200][i]
        x = r.uniform(0, 1, (n, 2))
        xs.append(x)
        ys.append(np.sin(3 * x.sum(1)) + s * r.normal(size=n))
        gps.append(G.gp_init(G.cf_sexp(lscale=0.5, magn=1.0), G.lik_gaussian(sigma=s)))
    B = GPBatch(gps, xs, ys, refine_iters=6)
    e, g = B.energy_grad(allow_not_pd=True)
    st, rr = B.status()
    assert st[1] in (L.GPL_ERR_NOCONV, L.GPL_ERR_NOT_PD) and e[1] == np.inf
    for i in (0, 2, 3):
        assert st[i] == L.GPL_OK and rr[i] < 1e-6
        fo = O.gp_fit([c.spec() for c in gps[i].cfs], gps[i].lik.spec(), xs[i], ys[i])
        assert e[i] == pytest.approx(O.gp_energy(fo), rel=1e-6)
    with pytest.raises(L.GplError) as ei:
        B.energy_grad()
    assert ei.value.code in (L.GPL_ERR_NOCONV, L.GPL_ERR_NOT_PD)
