"""Parity of the device path (through the C ABI) with the FP64 oracle.

Tolerance for the plain FP64 path: relative 1e-8 on the log evidence, its
gradient, the posterior mode and the predictive mean/variance (the stated
bar); covariance entries to 1e-12.
"""

import math
import os
import warnings

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-8
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# ----------------------------------------------------------- covariance builds

@pytest.mark.parametrize("n1,n2,d", [(1, 1, 1), (5, 7, 1), (63, 64, 3), (65, 130, 8), (200, 33, 16)])
def test_level0_cf_builders(gpl, n1, n2, d):
    from gplite_amd import _lib as L
    r = np.random.default_rng(n1 * 1000 + n2)
    x1 = r.uniform(-1, 1, (n1, d)); x2 = r.uniform(-1, 1, (n2, d))
    ls = r.uniform(0.2, 1.0, d)
    lib, ctx = L.lib(), L.ctx()
    for fn, fo in ((lib.gpl_cf_sexp, O.cf_sexp_c), (lib.gpl_cf_matern32, O.cf_matern32_c),
                   (lib.gpl_cf_matern52, O.cf_matern52_c)):
        for lsv in (ls, ls[:1]):
            out = np.empty((n1, n2), order="F")
            L.check(fn(ctx, L.dptr(L.f64c(x1)), n1, L.dptr(L.f64c(x2)), n2, d, L.dptr(np.ascontiguousarray(lsv)),
                       len(lsv), 1.3, 0, out.ctypes.data_as(L._dp)))
            ref = fo(x1, x2, lsv if len(lsv) > 1 else float(lsv[0]), 1.3)
            assert np.max(np.abs(out - ref)) <= 1e-12 * np.max(np.abs(ref))
            dg = np.empty(min(n1, n2))
            L.check(fn(ctx, L.dptr(L.f64c(x1)), n1, L.dptr(L.f64c(x2)), n2, d, L.dptr(np.ascontiguousarray(lsv)),
                       len(lsv), 1.3, 1, L.dptr(dg)))
            np.testing.assert_allclose(dg, fo(x1, x2, lsv if len(lsv) > 1 else float(lsv[0]), 1.3, diag_only=True),
                                       rtol=1e-12)
    out = np.empty((n1, n2), order="F")
    L.check(lib.gpl_cf_nn(ctx, L.dptr(L.f64c(x1)), n1, L.dptr(L.f64c(x2)), n2, d, 1.0, 3.0, 0.8, 0,
                          out.ctypes.data_as(L._dp)))
    np.testing.assert_allclose(out, O.cf_nn_c(x1, x2, 1.0, 3.0, 0.8), rtol=1e-11, atol=1e-14)


COMPOSITES = [
    lambda G: [G.cf_sexp(lscale=0.6), G.cf_lin(magn=0.4), G.cf_const(magn=0.3)],
    lambda G: [G.cf_prod(G.cf_sexp(lscale=0.9), G.cf_matern32(lscale=[1.1, 0.6], prior_magn=G.prior_fixed()))],
    lambda G: [G.cf_periodic(vars=[0], period=1.3, cf_base=G.cf_sexp(lscale=0.7, magn=1.2)),
               G.cf_matern52(vars=[1], lscale=0.5)],
    lambda G: [G.cf_nn(sigma0=1.0, sigma=2.0), G.cf_lin(normalize=True)],
    lambda G: [G.cf_periodic(period=2.0, cf_base=G.cf_matern32(lscale=[0.5, 0.6, 0.7, 0.8]))],
]


def _spec(cfs):
    return [c.spec() for c in cfs]


@pytest.mark.parametrize("ci", range(len(COMPOSITES)))
def test_eval_cf_composites(gpl, ci):
    G = gpl
    r = np.random.default_rng(ci)
    x1 = r.normal(size=(70, 2)); x2 = r.normal(size=(45, 2))
    cfs = COMPOSITES[ci](G)
    for c in cfs:
        c.learn_scales(x1)
    K = G.eval_cf(cfs, x1, x2)
    ref = O.eval_cf(_spec(cfs), x1, x2)
    assert rel(K, ref) < 1e-12
    np.testing.assert_allclose(G.eval_cf(cfs, x1, x2, diag_only=True), O.eval_cf(_spec(cfs), x1, x2, diag_only=True),
                               rtol=1e-12, atol=1e-15)


# ----------------------------------------------------------- Gaussian fit

@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 200, 333, 1000])
def test_gaussian_fit_ragged_sizes(gpl, n):
    G = gpl
    r = np.random.default_rng(n)
    x = r.normal(size=(n, 2)); y = np.sin(x.sum(1)) + 0.1 * r.normal(size=n)
    gp = G.gp_init(G.cf_matern52(lscale=[0.7, 1.3], magn=1.2), G.lik_gaussian(sigma=0.2))
    gf = G.gp_fit(gp, x, y)
    fo = O.gp_fit(_spec(gp.cfs), gp.lik.spec(), x, y)
    assert gf.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=RTOL)
    assert rel(gf.fit.alpha, fo["alpha"]) < RTOL
    assert rel(gf.fit.fmean, fo["fmean"]) < RTOL
    assert G.gp_energy(gf) == pytest.approx(O.gp_energy(fo), rel=RTOL)
    xt = r.normal(size=(37, 2))
    pr = G.gp_pred(gf, xt, var=True)
    mo, vo = O.gp_pred(fo, xt, var=True)
    assert rel(pr["mean"], mo) < RTOL
    assert rel(pr["var"], vo) < RTOL
    if n <= 200:
        np.testing.assert_allclose(gf.fit.C_chol, fo["C_chol"], rtol=1e-9, atol=1e-12)


def test_golden_c1(gpl):
    """The CPU-runnable config: n=200, cf_sexp(0.3, 1), sigma=0.1, pred at 300 points."""
    G = gpl
    ref = np.load(os.path.join(ROOT, "tests", "golden", "c1.npz"))
    gp = G.gp_init(G.cf_sexp(lscale=0.3, magn=1.0), G.lik_gaussian(sigma=0.1))
    gf = G.gp_fit(gp, ref["c1_x"], ref["c1_y"])
    assert gf.fit.log_evidence == pytest.approx(float(ref["c1_logz"]), rel=RTOL)
    assert G.gp_energy(gf) == pytest.approx(float(ref["c1_energy"]), rel=RTOL)
    assert rel(gf.fit.alpha, ref["c1_alpha"]) < RTOL
    pr = G.gp_pred(gf, ref["c1_xt"], var=True)
    assert rel(pr["mean"], ref["c1_mean"]) < RTOL
    assert rel(pr["var"], ref["c1_var"]) < RTOL
    assert rel(G.gp_energy_grad(gf), ref["c1_grad"]) < RTOL


def test_golden_classification(gpl):
    G = gpl
    ref = np.load(os.path.join(ROOT, "tests", "golden", "c1.npz"))
    gp = G.gp_init(G.cf_matern32(lscale=0.9, magn=1.5), G.lik_bernoulli("logit"),
                   approx=G.approx_laplace(maxiter=50, tol=1e-8))
    gf = G.gp_fit(gp, ref["cl_x"], ref["cl_y"])
    assert gf.fit.log_evidence == pytest.approx(float(ref["cl_la_logz"]), rel=RTOL)
    assert rel(gf.fit.fmean, ref["cl_la_fmean"]) < RTOL
    assert gf.fit.iters == int(ref["cl_la_iters"])
    gp = G.gp_init(G.cf_matern32(lscale=0.9, magn=1.5), G.lik_bernoulli("logit"), approx=G.approx_ep())
    ge = G.gp_fit(gp, ref["cl_x"], ref["cl_y"])
    assert ge.fit.log_evidence == pytest.approx(float(ref["cl_ep_logz"]), rel=RTOL)
    assert rel(ge.fit.fmean, ref["cl_ep_fmean"]) < RTOL
    assert rel(ge.fit.fvar, ref["cl_ep_fvar"]) < RTOL
    assert ge.fit.iters == int(ref["cl_ep_iters"])


GRAD_CASES = [
    lambda G: [G.cf_sexp(lscale=[0.7, 1.3], magn=1.2)],
    lambda G: [G.cf_matern32(lscale=0.9, magn=0.8)],
    lambda G: [G.cf_matern52(lscale=[0.5, 0.9], magn=1.1), G.cf_const(magn=0.5)],
    lambda G: [G.cf_nn(sigma0=1.0, sigma=1.5)],
    lambda G: [G.cf_lin(magn=0.6), G.cf_sexp(lscale=1.0)],
    lambda G: [G.cf_periodic(vars=[0], period=2.0, cf_base=G.cf_sexp(lscale=0.8))],
    lambda G: [G.cf_prod(G.cf_sexp(lscale=0.9), G.cf_matern32(lscale=[1.1, 0.6], magn=0.7, prior_magn=G.prior_fixed()))],
    # a product of a product (R/cf.R:189-215 allow nesting; it is one product)
    lambda G: [G.cf_prod(G.cf_prod(G.cf_sexp(vars=[0], lscale=0.9), G.cf_lin(vars=[1], prior_magn=G.prior_fixed())),
                         G.cf_matern52(lscale=1.2, prior_magn=G.prior_fixed())), G.cf_const(magn=0.3)],
]


@pytest.mark.parametrize("ci", range(len(GRAD_CASES)))
@pytest.mark.parametrize("n", [30, 150])
def test_energy_gradient(gpl, ci, n):
    G = gpl
    r = np.random.default_rng(ci + n)
    x = r.normal(size=(n, 2)); y = np.sin(x.sum(1)) + 0.1 * r.normal(size=n)
    gp = G.gp_init(GRAD_CASES[ci](G), G.lik_gaussian(sigma=0.25, prior_sigma=G.prior_lognormal(-1.0, 0.5)))
    gf = G.gp_fit(gp, x, y)
    fo = O.gp_fit(_spec(gf.cfs), gf.lik.spec(), x, y)
    g = G.gp_energy_grad(gf)
    go = O.gp_energy_grad(fo)
    assert rel(g, go) < RTOL, (g, go)


@pytest.mark.parametrize("n", [600, 1000, 2500])
def test_energy_gradient_inverse_split_sizes(gpl, n):
    """Padded sizes 640, 1024 and 2560: L^-1 is formed in pieces beside the
    Cholesky (the leading half, then the leading part of the back half and the
    top rows of their cross block, the rest after the last panel), with back
    halves of 128, 512 and 512 rows. The FP64 gradient matches the oracle at
    1e-8; the gradient of the mixed form (FP32 L^-1, B^-1) at 1e-3."""
    G = gpl
    r = np.random.default_rng(n)
    x = r.normal(size=(n, 3)); y = np.sin(x.sum(1)) + 0.1 * r.normal(size=n)
    gp = G.gp_init(G.cf_matern52(lscale=[0.9, 1.1, 1.3], magn=0.8), G.lik_gaussian(sigma=0.2))
    gf = G.gp_fit(gp, x, y)
    fo = O.gp_fit(_spec(gf.cfs), gf.lik.spec(), x, y)
    go = O.gp_energy_grad(fo)
    assert rel(G.gp_energy_grad(gf), go) < RTOL
    gm = G.gp_fit(gp, x, y, precision="mixed")
    assert rel(G.gp_energy_grad(gm), go) < 1e-3


@pytest.mark.parametrize("lik_name", ["logit", "probit", "poisson", "binomial"])
def test_laplace_parity(gpl, lik_name):
    G = gpl
    r = np.random.default_rng(11)
    n = 300
    x = r.normal(size=(n, 3))
    f = np.sin(2 * x[:, 0]) + 0.5 * x[:, 1]
    trials = None
    offset = None
    if lik_name == "poisson":
        lk = G.lik_poisson()
        offset = 0.2 * r.normal(size=n)
        y = r.poisson(np.exp(0.5 * f + offset)).astype(float)
    elif lik_name == "binomial":
        lk = G.lik_binomial("logit")
        trials = r.integers(1, 10, n).astype(float)
        y = r.binomial(trials.astype(int), 1 / (1 + np.exp(-f))).astype(float)
    else:
        lk = G.lik_bernoulli(lik_name)
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-f))).astype(float)
    gp = G.gp_init(G.cf_sexp(lscale=[0.8, 0.9, 1.2], magn=1.5), lk, approx=G.approx_laplace(maxiter=50, tol=1e-8))
    gf = G.gp_fit(gp, x, y, trials=trials, offset=offset)
    kw = {}
    if trials is not None:
        kw["trials"] = trials
    if offset is not None:
        kw["offset"] = offset
    fo = O.gp_fit(_spec(gf.cfs), lk.spec(), x, y, maxiter=50, tol=1e-8, **kw)
    assert gf.fit.iters == fo["iters"]
    assert gf.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=RTOL)
    assert rel(gf.fit.fmean, fo["fmean"]) < RTOL
    xt = r.normal(size=(50, 3))
    pr = G.gp_pred(gf, xt, var=True)
    mo, vo = O.gp_pred(fo, xt, var=True)
    assert rel(pr["mean"], mo) < RTOL and rel(pr["var"], vo) < RTOL


@pytest.mark.parametrize("lik_name", ["logit", "probit", "poisson"])
def test_ep_parity(gpl, lik_name):
    G = gpl
    r = np.random.default_rng(12)
    n = 200
    x = r.normal(size=(n, 2))
    f = np.sin(2 * x[:, 0]) + 0.5 * x[:, 1]
    if lik_name == "poisson":
        lk = G.lik_poisson()
        y = r.poisson(np.exp(0.5 * f)).astype(float)
    else:
        lk = G.lik_bernoulli(lik_name)
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-f))).astype(float)
    gp = G.gp_init(G.cf_sexp(lscale=0.8, magn=1.2), lk, approx=G.approx_ep())
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        gf = G.gp_fit(gp, x, y)
        fo = O.gp_fit(_spec(gf.cfs), lk.spec(), x, y, approx="ep")
    assert gf.fit.iters == fo["iters"]
    assert gf.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=RTOL)
    assert rel(gf.fit.fmean, fo["fmean"]) < RTOL
    assert rel(gf.fit.fvar, fo["fvar"]) < RTOL
    xt = r.normal(size=(20, 2))
    pr = G.gp_pred(gf, xt, var=True)
    mo, vo = O.gp_pred(fo, xt, var=True)
    assert rel(pr["mean"], mo) < RTOL and rel(pr["var"], vo) < RTOL


def test_prior_prediction_and_refit_guard(gpl):
    """R tests/testthat/test_pred.R:85-119: prior prediction unfitted; error after
    set_param until refit."""
    G = gpl
    r = np.random.default_rng(3)
    x = r.normal(size=(30, 1)); y = x[:, 0] ** 2
    gp = G.gp_init(G.cf_sexp(), G.lik_gaussian())
    xt = r.normal(size=(5, 1))
    pr = G.gp_pred(gp, xt, var=True)
    assert np.all(pr["mean"] == 0)
    np.testing.assert_allclose(pr["var"], O.eval_cf(_spec(gp.cfs), xt, xt, diag_only=True), rtol=1e-13)
    gf = G.gp_fit(gp, x, y)
    g2 = G.set_param(gf, np.array(list(G.get_param(gf).values())) + 0.1)
    with pytest.raises(RuntimeError):
        G.gp_pred(g2, xt)
    g3 = G.gp_fit(g2, x, y)
    G.gp_pred(g3, xt)


def test_closed_form_gaussian_suite(gpl):
    """R tests/testthat/test_basic.R:1-132 restated at 1e-8: every one of the seven
    covariance functions alone, every pair as a sum and every triple as a product
    (63 models), ten hyperparameter perturbations each, logZ and pred vs the
    closed form."""
    import copy
    import itertools
    G = gpl
    r = np.random.default_rng(1234)
    n, nt = 20, 30
    x = r.normal(size=n); xt = r.normal(size=nt); y = x ** 2 + 0.2 * r.normal(size=n)
    base = [G.cf_const(), G.cf_lin(), G.cf_sexp(), G.cf_matern32(), G.cf_matern52(), G.cf_nn(), G.cf_periodic()]
    cp = copy.deepcopy
    models = [[cp(c)] for c in base]
    models += [[cp(base[i]), cp(base[j])] for i, j in itertools.combinations(range(len(base)), 2)]
    with warnings.catch_warnings():  # products warn about several magnitudes (R's SWO)
        warnings.simplefilter("ignore")
        models += [[cp(base[i]) * cp(base[j]) * cp(base[k])]
                   for i, j, k in itertools.combinations(range(len(base)), 3)]
    assert len(models) == 63
    for cfs in models:
        gp = G.gp_init(cfs, G.lik_gaussian())
        p0 = np.array(list(G.get_param(gp).values()))
        for _ in range(10):
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                gpp = G.gp_fit(G.set_param(gp, p0 + 0.3 * r.normal(size=p0.size)), x, y)
            spec = _spec(gpp.cfs)
            K = O.eval_cf(spec, x[:, None], x[:, None])
            lz = O.gaussian_logz_closed_form(K, y, gpp.lik.sigma)
            assert gpp.fit.log_evidence == pytest.approx(lz, rel=RTOL)
            pr = G.gp_pred(gpp, xt, var=True)
            Kt = O.eval_cf(spec, xt[:, None], x[:, None])
            C = K + (gpp.lik.sigma ** 2 + O.JITTER) * np.eye(n)
            assert rel(pr["mean"], Kt @ np.linalg.solve(C, y)) < 1e-7
            vv = O.eval_cf(spec, xt[:, None], xt[:, None], diag_only=True) - np.sum(Kt.T * np.linalg.solve(C, Kt.T), 0)
            assert np.max(np.abs(pr["var"] - vv)) <= 1e-7 * max(1.0, np.max(np.abs(vv)))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_pred_ragged_test_sizes(gpl, prec):
    """gp_pred at m = 0, 1 and around the 64-row tile edges against the oracle:
    the mean to 1e-8 (FP64) or 1e-6 (mixed), the variance to 1e-8 or 1e-3,
    relative; m = 0 gives empty results."""
    G = gpl
    r = np.random.default_rng(77)
    n = 150
    x = r.normal(size=(n, 2)); y = np.sin(x.sum(1)) + 0.1 * r.normal(size=n)
    gp = G.gp_init(G.cf_matern32(lscale=[0.8, 1.1], magn=1.1), G.lik_gaussian(sigma=0.1))
    gf = G.gp_fit(gp, x, y, precision=prec)
    fo = O.gp_fit(_spec(gf.cfs), gf.lik.spec(), x, y)
    tm, tv = (1e-8, 1e-8) if prec == "fp64" else (1e-6, 1e-3)
    for m in (0, 1, 63, 64, 65, 129, 300):
        xt = r.normal(size=(m, 2))
        pr = G.gp_pred(gf, xt, var=True)
        assert pr["mean"].shape == (m,) and pr["var"].shape == (m,)
        if m == 0:
            continue
        mo, vo = O.gp_pred(fo, xt, var=True)
        assert rel(pr["mean"], mo) < tm, m
        assert rel(pr["var"], vo) < tv, m


@pytest.mark.parametrize("case", ["ep_binomial_trials", "ep_poisson_offset", "gaussian_offset"])
def test_trials_and_offset(gpl, case):
    """trials and offset through the device fits (R/lik.R:297-366,188-270):
    EP with binomial trials, EP with a Poisson offset, and the Gaussian
    likelihood with an offset, whose pseudo-data ignore it (z = y,
    R/lik.R:191) while its log-likelihood adds it (R/lik.R:302).  logZ, the
    posterior mean and the predictive mean/variance to 1e-8 of the oracle."""
    G = gpl
    r = np.random.default_rng(21)
    n = 150
    x = r.normal(size=(n, 2))
    f = np.sin(2 * x[:, 0]) + 0.5 * x[:, 1]
    kw = {}
    if case == "ep_binomial_trials":
        lk, ap = G.lik_binomial("logit"), G.approx_ep()
        kw["trials"] = r.integers(1, 8, n).astype(float)
        y = r.binomial(kw["trials"].astype(int), 1 / (1 + np.exp(-f))).astype(float)
    elif case == "ep_poisson_offset":
        lk, ap = G.lik_poisson(), G.approx_ep()
        kw["offset"] = 0.3 * r.normal(size=n)
        y = r.poisson(np.exp(0.5 * f + kw["offset"])).astype(float)
    else:
        lk, ap = G.lik_gaussian(sigma=0.3), G.approx_laplace()
        kw["offset"] = 0.5 + 0.2 * r.normal(size=n)
        y = f + kw["offset"] + 0.3 * r.normal(size=n)
    gp = G.gp_init(G.cf_sexp(lscale=[0.8, 1.0], magn=1.2), lk, approx=ap)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        gf = G.gp_fit(gp, x, y, **kw)
        fo = O.gp_fit(_spec(gf.cfs), lk.spec(), x, y, approx="ep" if case.startswith("ep") else "laplace", **kw)
    assert gf.fit.iters == fo["iters"]
    assert gf.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=RTOL)
    assert rel(gf.fit.fmean, fo["fmean"]) < RTOL
    xt = r.normal(size=(40, 2))
    pr = G.gp_pred(gf, xt, var=True)
    mo, vo = O.gp_pred(fo, xt, var=True)
    assert rel(pr["mean"], mo) < RTOL and rel(pr["var"], vo) < RTOL
