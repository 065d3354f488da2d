"""Pins the FP64 oracle against mathematics that is independent of it:
library routines (scipy, sklearn, numpy), closed forms, Monte Carlo,
numerical integration and finite differences."""

import math
import warnings

import numpy as np
import pytest
import scipy.integrate as si
import scipy.linalg as sla
import scipy.stats as st
from sklearn.gaussian_process import kernels as skk

import oracle as O


def _rng(s=0):
    return np.random.default_rng(s)


# ---------------------------------------------------------------- covariances

@pytest.mark.parametrize("d", [1, 3, 8])
def test_stationary_vs_sklearn_ard(d):
    r = _rng(d)
    x1 = r.uniform(-1, 1, (23, d)); x2 = r.uniform(-1, 1, (17, d))
    ls = r.uniform(0.2, 1.0, d)
    m = 1.7
    np.testing.assert_allclose(O.cf_sexp_c(x1, x2, ls, m),
                               m * m * skk.RBF(ls)(x1, x2), rtol=1e-13)
    np.testing.assert_allclose(O.cf_matern32_c(x1, x2, ls, m),
                               m * m * skk.Matern(ls, nu=1.5)(x1, x2), rtol=1e-12)
    np.testing.assert_allclose(O.cf_matern52_c(x1, x2, ls, m),
                               m * m * skk.Matern(ls, nu=2.5)(x1, x2), rtol=1e-12)
    # isotropic scalar lscale == ARD with equal entries
    np.testing.assert_allclose(O.cf_matern52_c(x1, x2, 0.4, m),
                               O.cf_matern52_c(x1, x2, np.full(d, 0.4), m), rtol=1e-14)


def test_stationary_special_values():
    x = np.array([[0.0]]); y = np.array([[0.3]])
    assert O.cf_sexp_c(x, x, 0.3, 2.0)[0, 0] == 4.0
    assert O.cf_sexp_c(x, y, 0.3, 1.0)[0, 0] == pytest.approx(math.exp(-0.5), rel=1e-15)
    assert O.cf_matern32_c(x, y, 0.3, 1.0)[0, 0] == pytest.approx((1 + math.sqrt(3)) * math.exp(-math.sqrt(3)), rel=1e-14)
    assert O.cf_matern52_c(x, y, 0.3, 1.0)[0, 0] == pytest.approx((1 + math.sqrt(5) + 5 / 3) * math.exp(-math.sqrt(5)), rel=1e-14)


def test_nn_kernel_monte_carlo():
    """Williams (1998): k = E_u[erf(u.xt) erf(u.xt')], u ~ N(0, s^2 diag(s0^2, 1..1))."""
    from scipy.special import erf
    r = _rng(5)
    d = 2
    x = r.normal(size=(3, d))
    s0, s, m = 1.0, 0.8, 1.0
    K = O.cf_nn_c(x, x, s0, s, m)
    N = 2_000_000
    u0 = r.normal(size=N) * s0 * s
    u = r.normal(size=(N, d)) * s
    h = erf(u0[:, None] + u @ x.T)
    Kmc = (h.T @ h) / N
    np.testing.assert_allclose(K, Kmc, atol=4e-3)


def test_periodic_closed_form():
    """cos/sin feature map + sexp == MacKay's exp(-2 sin^2(pi dx/p)/l^2)."""
    r = _rng(2)
    x1 = r.uniform(-3, 3, (11, 1)); x2 = r.uniform(-3, 3, (9, 1))
    p, l, m = 1.3, 0.7, 1.2
    cf = {"kind": "periodic", "period": p, "base": {"kind": "sexp", "lscale": l, "magn": m}}
    K = O.eval_cf([cf], x1, x2)
    dx = x1 - x2.T
    Kc = m * m * np.exp(-2 * np.sin(math.pi * dx / p) ** 2 / l ** 2)
    np.testing.assert_allclose(K, Kc, rtol=1e-12)


def test_composition_and_diag():
    r = _rng(3)
    x = r.normal(size=(12, 2))
    a = {"kind": "sexp", "lscale": 0.5, "magn": 1.1}
    b = {"kind": "lin", "magn": 0.7}
    c = {"kind": "const", "magn": 0.4}
    Ka = O.eval_cf([a], x, x); Kb = O.eval_cf([b], x, x); Kc = O.eval_cf([c], x, x)
    np.testing.assert_allclose(O.eval_cf([a, b, c], x, x), Ka + Kb + Kc, rtol=1e-14)
    np.testing.assert_allclose(O.eval_cf([{"kind": "prod", "cfs": [a, b, c]}], x, x), Ka * Kb * Kc, rtol=1e-14)
    np.testing.assert_allclose(Kb, 0.49 * x @ x.T, rtol=1e-14)
    for cf in (a, b, c, {"kind": "nn", "sigma0": 1.0, "sigma": 3.0, "magn": 1.0},
               {"kind": "matern32", "lscale": 0.3, "magn": 1.0},
               {"kind": "periodic", "period": 1.0, "base": {"kind": "matern52", "lscale": 0.3, "magn": 1.0}}):
        K = O.eval_cf([cf], x, x)
        np.testing.assert_allclose(O.eval_cf([cf], x, x, diag_only=True), np.diag(K), rtol=1e-14)
        np.testing.assert_allclose(K, K.T, rtol=1e-14)


def test_normalization_invariance():
    """R tests/testthat/test_normalization.R:81-139: normalize=TRUE on raw x
    equals the plain kernel on pre-scaled x."""
    r = _rng(4)
    x = 5 * (r.normal(size=(20, 1)) - 1)
    xs = (x - x.mean(0)) / x.std(0, ddof=1)
    for kind in ("sexp", "matern32", "matern52", "lin", "nn"):
        cf = {"kind": kind, "lscale": 0.3, "magn": 1.0, "sigma0": 1.0, "sigma": 3.0}
        cfn = dict(cf, normalize=True)
        cfp = dict(cf, normalize=False)
        K1 = O.eval_cf(O.learn_scales([cfn], x), x, x)
        K2 = O.eval_cf([cfp], xs, xs)
        np.testing.assert_allclose(K1, K2, rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------------- likelihoods

def test_loglik_vs_scipy():
    r = _rng(6)
    f = r.normal(size=50)
    y = r.integers(0, 2, 50).astype(float)
    ll = O.loglik({"kind": "bernoulli", "link": "logit"}, f, y, total=False)
    np.testing.assert_allclose(ll, st.bernoulli.logpmf(y, 1 / (1 + np.exp(-f))), rtol=1e-12)
    ll = O.loglik({"kind": "bernoulli", "link": "probit"}, f, y, total=False)
    np.testing.assert_allclose(ll, st.bernoulli.logpmf(y, st.norm.cdf(f)), rtol=1e-12)
    t = r.integers(1, 10, 50).astype(float)
    yb = np.minimum(r.integers(0, 10, 50), t)
    ll = O.loglik({"kind": "binomial", "link": "logit"}, f, yb, trials=t, total=False)
    np.testing.assert_allclose(ll, st.binom.logpmf(yb, t, 1 / (1 + np.exp(-f))), rtol=1e-11)
    yp = r.poisson(2.0, 50).astype(float)
    ll = O.loglik({"kind": "poisson", "link": "log"}, f, yp, total=False)
    np.testing.assert_allclose(ll, st.poisson.logpmf(yp, np.exp(f)), rtol=1e-12)
    ll = O.loglik({"kind": "gaussian", "sigma": 0.3}, f, y, total=False)
    np.testing.assert_allclose(ll, st.norm.logpdf(y, f, 0.3), rtol=1e-13)


@pytest.mark.parametrize("lik", [{"kind": "bernoulli", "link": "logit"}, {"kind": "bernoulli", "link": "probit"},
                                 {"kind": "poisson", "link": "log"}])
def test_loglik_derivative_fd(lik):
    r = _rng(7)
    f = r.normal(size=20)
    y = (r.integers(0, 2, 20) if lik["kind"] == "bernoulli" else r.poisson(2.0, 20)).astype(float)
    h = 1e-6
    fd = (O.loglik(lik, f + h, y, total=False) - O.loglik(lik, f - h, y, total=False)) / (2 * h)
    np.testing.assert_allclose(O.loglik_d(lik, f, y), fd, rtol=1e-6, atol=1e-8)
    g, g2 = O.loglik_d2(lik, f, y)
    fd2 = (O.loglik_d(lik, f + h, y) - O.loglik_d(lik, f - h, y)) / (2 * h)
    np.testing.assert_allclose(g2, fd2, rtol=1e-4, atol=1e-6)


# ---------------------------------------------------------------- quadrature

@pytest.mark.parametrize("p", [3, 7, 11, 15])
def test_gauss_hermite_vs_numpy(p):
    x, w = O.gauss_hermite_points(p)
    xr, wr = np.polynomial.hermite_e.hermegauss(p)
    np.testing.assert_allclose(x, xr, atol=1e-10)
    np.testing.assert_allclose(w, wr / math.sqrt(2 * math.pi), rtol=1e-8)
    assert abs(w.sum() - 1) < 1e-12


# ---------------------------------------------------------------- Gaussian fit

def _toy(n=40, d=2, seed=8):
    r = _rng(seed)
    x = r.normal(size=(n, d)); y = np.sin(x.sum(1)) + 0.1 * r.normal(size=n)
    return x, y


def test_gaussian_logz_closed_form_and_scipy():
    x, y = _toy()
    cfs = [{"kind": "matern52", "lscale": [0.7, 1.3], "magn": 1.2}]
    lik = {"kind": "gaussian", "sigma": 0.2}
    fit = O.gp_fit(cfs, lik, x, y)
    Kcf = O.eval_cf(cfs, x, x)
    ref = st.multivariate_normal(np.zeros(len(y)), Kcf + (0.04 + O.JITTER) * np.eye(len(y))).logpdf(y)
    assert fit["log_evidence"] == pytest.approx(ref, rel=1e-12)
    assert O.gaussian_logz_closed_form(Kcf, y, 0.2) == pytest.approx(ref, rel=1e-12)
    Ky = Kcf + (0.04 + O.JITTER) * np.eye(len(y))
    np.testing.assert_allclose(fit["alpha"], np.linalg.solve(Ky, y), rtol=1e-10)
    np.testing.assert_allclose(fit["C_chol"] @ fit["C_chol"].T, Ky, rtol=1e-12, atol=1e-12)


def test_pred_closed_form():
    """R tests/testthat/test_basic.R:99-132 restated at 1e-10."""
    x, y = _toy()
    xt = _rng(9).normal(size=(15, 2))
    cfs = [{"kind": "sexp", "lscale": 0.8, "magn": 1.0}, {"kind": "lin", "magn": 0.3}]
    lik = {"kind": "gaussian", "sigma": 0.3}
    fit = O.gp_fit(cfs, lik, x, y)
    mean, var = O.gp_pred(fit, xt, var=True)
    K = O.eval_cf(cfs, x, x) + O.JITTER * np.eye(len(y)); Kt = O.eval_cf(cfs, xt, x); Ktt = O.eval_cf(cfs, xt, xt)
    C = K + 0.09 * np.eye(len(y))
    np.testing.assert_allclose(mean, Kt @ np.linalg.solve(C, y), rtol=1e-10)
    np.testing.assert_allclose(var, np.diag(Ktt - Kt @ np.linalg.solve(C, Kt.T)), rtol=1e-9)
    m0, v0 = O.gp_pred(fit, xt, var=True, prior=True)
    assert np.all(m0 == 0); np.testing.assert_allclose(v0, np.diag(Ktt), rtol=1e-14)


ALL_CFS = [
    [{"kind": "sexp", "lscale": [0.7, 1.3], "magn": 1.2}],
    [{"kind": "matern32", "lscale": 0.9, "magn": 0.8}],
    [{"kind": "matern52", "lscale": [0.5, 0.9], "magn": 1.1}, {"kind": "const", "magn": 0.5}],
    [{"kind": "nn", "sigma0": 1.0, "sigma": 1.5, "magn": 1.0}],
    [{"kind": "lin", "magn": 0.6}, {"kind": "sexp", "lscale": 1.0, "magn": 1.0}],
    [{"kind": "periodic", "period": 2.0, "vars": [0], "base": {"kind": "sexp", "lscale": 0.8, "magn": 1.0}}],
    [{"kind": "periodic", "period": 2.0, "vars": [1], "base": {"kind": "matern52", "lscale": 0.8, "magn": 1.0}}],
    [{"kind": "prod", "cfs": [{"kind": "sexp", "lscale": 0.9, "magn": 1.0},
                               {"kind": "matern32", "lscale": [1.1, 0.6], "magn": 0.7, "priors": {"magn": {"kind": "fixed"}}}]}],
]


@pytest.mark.parametrize("ci", range(len(ALL_CFS)))
def test_energy_gradient_finite_difference(ci):
    """Analytic gradient (R&W eq 5.9) vs central differences of the energy."""
    x, y = _toy(n=30)
    cfs = ALL_CFS[ci]
    lik = {"kind": "gaussian", "sigma": 0.25, "priors": {"sigma": {"kind": "lognormal", "loc": -1.0, "scale": 0.5}}}
    fit = O.gp_fit(cfs, lik, x, y)
    g = O.gp_energy_grad(fit)
    p0 = O.get_param(cfs, lik)
    assert len(g) == len(p0) == len(O.param_names(cfs, lik))
    h = 1e-5
    for j in range(len(p0)):
        e = []
        for s in (+1, -1):
            p = p0.copy(); p[j] += s * h
            c2, l2 = O.set_param(cfs, lik, p)
            e.append(O.gp_energy(O.gp_fit(c2, l2, x, y)))
        fd = (e[0] - e[1]) / (2 * h)
        assert g[j] == pytest.approx(fd, rel=2e-6, abs=2e-7), (ci, j)


def test_param_roundtrip():
    """R tests/testthat/test_param.R:73-84: get_param(set_param(p)) == p."""
    cfs = ALL_CFS[7] + ALL_CFS[5]
    lik = {"kind": "gaussian", "sigma": 0.5}
    p = _rng(1).normal(size=len(O.get_param(cfs, lik)))
    c2, l2 = O.set_param(cfs, lik, p)
    np.testing.assert_allclose(O.get_param(c2, l2), p, rtol=1e-14)


def test_priors_vs_scipy():
    p = 0.37
    cf = {"kind": "nn", "sigma0": math.exp(p), "sigma": 1.0, "magn": 1.0,
          "priors": {"sigma": {"kind": "fixed"}, "magn": {"kind": "fixed"},
                     "sigma0": {"kind": "half_t", "df": 3.0, "scale": 2.0}}}
    lp = O.lpdf_prior([cf], {"kind": "bernoulli"})
    th = math.exp(p)
    ref = math.log(2) - math.log(2.0) + st.t.logpdf(th / 2.0, 3.0) + p
    assert lp == pytest.approx(ref, rel=1e-12)
    lik = {"kind": "gaussian", "sigma": math.exp(p), "priors": {"sigma": {"kind": "lognormal", "loc": 0.2, "scale": 0.7}}}
    assert O.lpdf_prior([], lik) == pytest.approx(st.norm.logpdf(p, 0.2, 0.7), rel=1e-12)
    for obj in ([cf], []):
        lk = lik if obj == [] else {"kind": "bernoulli"}
        g = O.lpdf_prior_grad(obj, lk)[0]
        h = 1e-6
        pp = O.get_param(obj, lk)
        fd = (O.lpdf_prior(*O.set_param(obj, lk, pp + h)) - O.lpdf_prior(*O.set_param(obj, lk, pp - h))) / (2 * h)
        assert g == pytest.approx(fd, rel=1e-6)


# ---------------------------------------------------------------- Laplace / EP

def _class_data(n=60, seed=10, kind="bernoulli"):
    r = _rng(seed)
    x = r.normal(size=(n, 2))
    K = O.eval_cf([{"kind": "sexp", "lscale": 0.8, "magn": 2.0}], x, x) + 1e-8 * np.eye(n)
    f = sla.cholesky(K, lower=True) @ r.normal(size=n)
    if kind == "bernoulli":
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-f))).astype(float)
    else:
        y = r.poisson(np.exp(0.5 * f)).astype(float)
    return x, y


@pytest.mark.parametrize("lik", [{"kind": "bernoulli", "link": "logit"}, {"kind": "bernoulli", "link": "probit"},
                                 {"kind": "poisson", "link": "log"}])
def test_laplace_mode_fixed_point(lik):
    """At the mode f = K grad log p(y|f) (R&W eq 3.17), independent of W."""
    x, y = _class_data(kind="poisson" if lik["kind"] == "poisson" else "bernoulli")
    cfs = [{"kind": "sexp", "lscale": 0.8, "magn": 2.0 if lik["kind"] != "poisson" else 0.7}]
    fit = O.gp_fit(cfs, lik, x, y, maxiter=100, tol=1e-10)
    f = fit["fmean"]
    np.testing.assert_allclose(f, fit["K"] @ O.loglik_d(lik, f, y), atol=1e-8)
    # logZ vs R&W Alg 3.2 with the analytic W at the mode (FD curvature differs ~1e-6)
    K = fit["K"]
    h = 1e-5
    W = -(O.loglik_d(lik, f + h, y) - O.loglik_d(lik, f - h, y)) / (2 * h)
    sW = np.sqrt(W)
    L = sla.cholesky(np.eye(len(y)) + sW[:, None] * K * sW[None, :], lower=True)
    a = O.loglik_d(lik, f, y)
    lz = -0.5 * a @ f + O.loglik(lik, f, y) - np.sum(np.log(np.diag(L)))
    assert fit["log_evidence"] == pytest.approx(lz, rel=1e-6)


@pytest.mark.parametrize("lik", [{"kind": "bernoulli", "link": "logit"}, {"kind": "poisson", "link": "log"}])
def test_laplace_evidence_uses_the_forward_difference_curvature(lik):
    """The reference's W is a forward difference of the analytic gradient with
    eps = 1e-6 (R/lik.R:449-454): W_fd = W + dW with dW_i = -(eps/2) g'''(f_i) to
    first order.  At the mode (W-independent, converged to 1e-11 so that W of the
    previous iterate equals W at the mode), the evidence then differs from the one
    with the analytic W by -1/2 [log|B(W_fd)| - log|B(W)|] = -1/2 sum_i S_ii dW_i,
    S = (K^-1 + W)^-1 the Laplace posterior covariance.  The observed gap must
    match that prediction, which is far from zero: an oracle using the analytic
    W, a central difference, or W of the new iterate fails here."""
    x, y = _class_data(n=80, seed=21, kind="poisson" if lik["kind"] == "poisson" else "bernoulli")
    cfs = [{"kind": "sexp", "lscale": 0.8, "magn": 2.0 if lik["kind"] != "poisson" else 0.7}]
    fit = O.gp_fit(cfs, lik, x, y, maxiter=200, tol=1e-11)
    f, K, n = fit["fmean"], fit["K"], len(y)
    # analytic derivatives of log p(y|f) in f: logit g'' = -mu(1-mu), g''' = -mu(1-mu)(1-2mu);
    # log link g'' = g''' = -exp(f)
    if lik["kind"] == "poisson":
        g2 = -np.exp(f)
        g3 = -np.exp(f)
    else:
        mu = 1 / (1 + np.exp(-f))
        g2 = -mu * (1 - mu)
        g3 = -mu * (1 - mu) * (1 - 2 * mu)
    W = -g2
    sW = np.sqrt(W)
    Lb = sla.cholesky(np.eye(n) + sW[:, None] * K * sW[None, :], lower=True)
    a = O.loglik_d(lik, f, y)
    lz_analytic = O.loglik(lik, f, y) - 0.5 * a @ f - np.sum(np.log(np.diag(Lb)))
    S = K - K @ (sW[:, None] * sla.cho_solve((Lb, True), sW[:, None] * K))
    dW = -0.5e-6 * g3
    predicted = -0.5 * np.sum(np.diag(S) * dW)
    observed = fit["log_evidence"] - lz_analytic
    assert abs(predicted) > 1e-8
    assert observed == pytest.approx(predicted, rel=2e-2, abs=1e-11)


_LATENT_LIKS = [({"kind": "bernoulli", "link": "logit"}, "bernoulli"), ({"kind": "bernoulli", "link": "probit"}, "bernoulli"),
                ({"kind": "poisson", "link": "log"}, "poisson"), ({"kind": "binomial", "link": "logit"}, "binomial"),
                ({"kind": "binomial", "link": "probit"}, "binomial")]


def _latent_data(kind, n=40, seed=0):
    r = _rng(seed)
    x = r.uniform(-2, 2, (n, 2))
    f = np.sin(x[:, 0]) + 0.5 * x[:, 1]
    tr = None
    if kind == "bernoulli":
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-2 * f))).astype(float)
    elif kind == "poisson":
        y = r.poisson(np.exp(f)).astype(float)
    else:
        tr = r.integers(1, 8, n).astype(float)
        y = r.binomial(tr.astype(int), 1 / (1 + np.exp(-f))).astype(float)
    return x, y, tr


@pytest.mark.parametrize("lik,kind", _LATENT_LIKS)
def test_exact_second_and_third_derivatives_match_central_differences(lik, kind):
    """loglik_d2_exact and loglik_d3 against central differences of loglik_d and
    loglik_d2_exact (step 1e-5, error ~1e-10), over f in [-4, 4] and the data's y
    (beyond, the reference's probit gradient forms 1 - Phi(f) by subtraction and its
    differences lose the digits the exact forms keep)."""
    x, y, tr = _latent_data(kind, n=200, seed=3)
    f = np.linspace(-4, 4, len(y))
    h = 1e-5
    d2 = O.loglik_d2_exact(lik, f, y, trials=tr)
    d2c = (O.loglik_d(lik, f + h, y, trials=tr) - O.loglik_d(lik, f - h, y, trials=tr)) / (2 * h)
    np.testing.assert_allclose(d2, d2c, rtol=1e-6, atol=1e-9)
    d3 = O.loglik_d3(lik, f, y, trials=tr)
    d3c = (O.loglik_d2_exact(lik, f + h, y, trials=tr) - O.loglik_d2_exact(lik, f - h, y, trials=tr)) / (2 * h)
    np.testing.assert_allclose(d3, d3c, rtol=1e-5, atol=1e-8)


def _grad_fd(cfs_of, th, fit_of, h=1e-5):
    fd = []
    for j in range(len(th)):
        tp = th.copy(); tp[j] *= np.exp(h)
        tm = th.copy(); tm[j] *= np.exp(-h)
        fd.append(-(fit_of(tp)["log_evidence"] - fit_of(tm)["log_evidence"]) / (2 * h))
    return np.array(fd)


@pytest.mark.parametrize("lik,kind", _LATENT_LIKS)
def test_laplace_gradient_with_exact_curvature_matches_finite_differences(lik, kind):
    """R&W Algorithm 5.1 (explicit term plus the mode's implicit dependence through
    d3 log p) against central differences of the Laplace evidence with the exact
    curvature, Newton run to 1e-13: a dropped implicit term or a wrong sign of it
    moves the result by ~10%."""
    x, y, tr = _latent_data(kind)
    cfs_of = lambda t: [{"kind": "sexp", "lscale": np.array(t[:2]), "magn": t[2]}]
    fit_of = lambda t: O.gp_fit(cfs_of(t), lik, x, y, maxiter=200, tol=1e-13, analytic_w=True, trials=tr)
    th = np.array([0.8, 1.3, 1.1])
    g = O.gp_energy_grad(fit_of(th), include_prior=False)
    fd = _grad_fd(cfs_of, th, fit_of)
    np.testing.assert_allclose(g, fd, rtol=1e-7, atol=1e-8 * np.max(np.abs(fd)))
    # the same evidence by R&W Algorithm 3.2 at the mode
    fit = fit_of(th)
    f, K = fit["fmean"], fit["K"]
    sW = np.sqrt(-O.loglik_d2_exact(lik, f, y, trials=tr))
    L = sla.cholesky(np.eye(len(y)) + sW[:, None] * K * sW[None, :], lower=True)
    a = O.loglik_d(lik, f, y, trials=tr)
    lz = -0.5 * a @ f + O.loglik(lik, f, y, trials=tr) - np.sum(np.log(np.diag(L)))
    assert fit["log_evidence"] == pytest.approx(lz, rel=1e-12)


@pytest.mark.parametrize("lik,kind,fine,coarse_tol", [(_LATENT_LIKS[0] + (31, 1e-4)), (_LATENT_LIKS[1] + (31, 1e-4)),
                                                       (_LATENT_LIKS[3] + (31, 1e-3)), (_LATENT_LIKS[2] + (61, 5e-2))])
def test_ep_gradient_matches_finite_differences(lik, kind, fine, coarse_tol):
    """The EP gradient at the fixed point (site parameters held: R&W eq. 5.27)
    against central differences of log Z_EP with EP run to 1e-12.  With a fine
    Gauss-Hermite rule (31 nodes; 61 for Poisson counts up to 8, whose tilted
    densities are narrow against the cavity) the tilted moments are exact and the
    two agree to 1e-6.  With the reference's 11 nodes the moments and log Z_site
    disagree at the quadrature error: 1e-5 to 1e-4 for the binary and binomial data, 2.5%
    for the counts."""
    x, y, tr = _latent_data(kind)
    cfs_of = lambda t: [{"kind": "sexp", "lscale": np.array(t[:2]), "magn": t[2]}]
    th = np.array([0.8, 1.3, 1.1])
    for qo, tol in ((fine, 1e-6), (11, coarse_tol)):
        fit_of = lambda t: O.gp_fit(cfs_of(t), lik, x, y, approx="ep", maxiter=1000, ep_tol=1e-12, quad_order=qo,
                                    trials=tr)
        g = O.gp_energy_grad(fit_of(th), include_prior=False)
        fd = _grad_fd(cfs_of, th, fit_of)
        np.testing.assert_allclose(g, fd, rtol=tol, atol=tol * np.max(np.abs(fd)))


def test_ep_single_site_is_exact():
    """n=1: the EP evidence equals log int N(f|0,k) p(y|f) df."""
    for lik, y in (({"kind": "bernoulli", "link": "probit"}, 1.0), ({"kind": "bernoulli", "link": "logit"}, 0.0),
                   ({"kind": "poisson", "link": "log"}, 3.0)):
        k = 1.7
        # with 11 nodes the evidence is the 11-point probabilists' GH rule (numpy's nodes)
        fit = O.ep(lik, np.array([[k]]), np.array([y]), maxiter=200, quad_order=11)
        xg, wg = np.polynomial.hermite_e.hermegauss(11)
        gh = np.sum(wg / math.sqrt(2 * math.pi) * np.exp(O.loglik(lik, math.sqrt(k) * xg, np.full(11, y), total=False)))
        assert fit["log_evidence"] == pytest.approx(math.log(gh), abs=1e-10)
        # with many nodes it converges to the exact integral
        fit = O.ep(lik, np.array([[k]]), np.array([y]), maxiter=200, quad_order=61)
        integ = si.quad(lambda f: st.norm.pdf(f, 0, math.sqrt(k)) * math.exp(O.loglik(lik, np.array([f]), np.array([y]))),
                        -30, 30, epsabs=1e-13)[0]
        assert fit["log_evidence"] == pytest.approx(math.log(integ), abs=1e-5)


def test_ep_fixed_point_moment_matching():
    x, y = _class_data(n=40)
    lik = {"kind": "bernoulli", "link": "probit"}
    cfs = [{"kind": "sexp", "lscale": 0.8, "magn": 2.0}]
    x = x; K = O.eval_cf(cfs, x, x) + 1e-6 * np.eye(len(y))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        fit = O.ep(lik, K, y, maxiter=500, tol=1e-10, damping=0.9)
    # posterior marginals = tilted moments of the cavity x likelihood
    mt, vt, *_ = O.tilted_moments(lik, fit["fmean"], 1 / fit["fvar"], fit["z"], fit["P"], y, 11)
    np.testing.assert_allclose(fit["fmean"], mt, atol=1e-7)
    np.testing.assert_allclose(fit["fvar"], vt, rtol=1e-6)
    # fvar is the diagonal of (K^-1 + diag(P))^-1
    S = np.linalg.inv(np.linalg.inv(K) + np.diag(fit["P"]))
    np.testing.assert_allclose(fit["fvar"], np.diag(S), rtol=1e-8)
    np.testing.assert_allclose(fit["fmean"], S @ (fit["P"] * fit["z"]), rtol=1e-7, atol=1e-9)


def test_cpu_baseline_matches_oracle():
    """The LAPACK-based CPU baseline computes the same evidence and gradient."""
    from oracle.cpu_baseline import energy_grad_matern52_ard
    r = _rng(21)
    x = r.uniform(0, 1, (90, 3)); y = np.sin(x.sum(1)) + 0.05 * r.normal(size=90)
    ls = np.array([0.3, 0.5, 0.8])
    lz, g, _ = energy_grad_matern52_ard(x, y, ls, 1.1, 0.07)
    cfs = [{"kind": "matern52", "lscale": ls, "magn": 1.1}]
    fit = O.gp_fit(cfs, {"kind": "gaussian", "sigma": 0.07}, x, y)
    assert lz == pytest.approx(fit["log_evidence"], rel=1e-10)
    np.testing.assert_allclose(g, O.gp_energy_grad(fit), rtol=1e-8)


def test_oracle_draw_gaussian_closed_form():
    """Joint posterior draws (R/gp_draw.R:101-118).
    * Gaussian likelihood: mean = K_t K_y^-1 y and cov = K_tt - K_t K_y^-1 K_t^T
      + jitter I (Rasmussen & Williams eqs. 2.23-2.24), from a direct solve with
      K_y = K_cf + (sigma^2 + jitter) I.
    * With z = I the draws minus the mean are the factor itself: lower
      triangular, positive diagonal and L L^T = cov, which fixes L uniquely.
    * diag(cov) - jitter equals the predictive variance of gp_pred."""
    r = np.random.default_rng(11)
    n, m = 50, 17
    x = r.normal(size=(n, 2))
    y = np.cos(x.sum(1)) + 0.1 * r.normal(size=n)
    xt = r.normal(size=(m, 2))
    cf = [{"kind": "matern52", "lscale": 0.8, "magn": 1.1}]
    lik = {"kind": "gaussian", "sigma": 0.3}
    fit = O.gp_fit(cf, lik, x, y)
    jit = 1e-6
    Kc = O.eval_cf(fit["cfs"], x, x)
    Kt = O.eval_cf(fit["cfs"], xt, x)
    Ky = Kc + (0.3 ** 2 + jit) * np.eye(n)
    mean_cf = Kt @ np.linalg.solve(Ky, y)
    cov_cf = O.eval_cf(fit["cfs"], xt, xt) - Kt @ np.linalg.solve(Ky, Kt.T) + jit * np.eye(m)
    draws = O.gp_draw(fit, xt, np.eye(m), jitter=jit)
    Lf = draws - mean_cf[:, None]
    assert np.max(np.abs(np.triu(Lf, 1))) < 1e-12
    assert np.all(np.diag(Lf) > 0)
    np.testing.assert_allclose(Lf @ Lf.T, cov_cf, atol=1e-10)
    mean, cov = O.gp_pred_cov(fit, xt, jitter=jit)
    np.testing.assert_allclose(mean, mean_cf, atol=1e-10)
    _, v = O.gp_pred(fit, xt, var=True)
    np.testing.assert_allclose(np.diag(cov) - jit, v, atol=1e-12)


def test_oracle_pred_cfind_closed_form():
    """cfind (R/gp_pred.R:206-207): for an additive Gaussian model the posterior
    of one component f_k at xt is mean K_k,t K_y^-1 y and variance
    diag(K_k,tt - K_k,t K_y^-1 K_k,t^T) (a direct solve against the full K_y);
    the component means sum to the full mean; cfind over every function equals
    no cfind."""
    r = np.random.default_rng(12)
    n, m = 80, 25
    x = r.normal(size=(n, 2))
    y = np.sin(2 * x[:, 0]) + 0.5 * x[:, 1] + 0.2 * r.normal(size=n)
    xt = r.normal(size=(m, 2))
    cfs = [{"kind": "sexp", "lscale": 0.7, "magn": 1.0, "vars": [0]},
           {"kind": "matern32", "lscale": 1.5, "magn": 0.8, "vars": [1]},
           {"kind": "const", "magn": 0.4}]
    lik = {"kind": "gaussian", "sigma": 0.2}
    fit = O.gp_fit(cfs, lik, x, y)
    c2 = fit["cfs"]
    Ky = O.eval_cf(c2, x, x) + (0.2 ** 2 + 1e-6) * np.eye(n)
    means = []
    for k in range(3):
        Kt = O.eval_cf(c2, xt, x, cfind=[k])
        mu = Kt @ np.linalg.solve(Ky, y)
        v = np.diag(O.eval_cf(c2, xt, xt, cfind=[k]) - Kt @ np.linalg.solve(Ky, Kt.T))
        mo, vo = O.gp_pred(fit, xt, var=True, cfind=[k])
        np.testing.assert_allclose(mo, mu, atol=1e-10)
        np.testing.assert_allclose(vo, v, atol=1e-10)
        means.append(mo)
    mfull, vfull = O.gp_pred(fit, xt, var=True)
    np.testing.assert_allclose(sum(means), mfull, atol=1e-10)
    ma, va = O.gp_pred(fit, xt, var=True, cfind=[0, 1, 2])
    np.testing.assert_allclose(ma, mfull, atol=1e-12)
    np.testing.assert_allclose(va, vfull, atol=1e-12)


def _fitc_case(seed=14, n=90, m=12, kind="gaussian"):
    r = np.random.default_rng(seed)
    x = r.normal(size=(n, 2))
    f = np.sin(x.sum(1))
    z = r.normal(size=(m, 2)) * 1.2
    if kind == "gaussian":
        y = f + 0.3 * r.normal(size=n)
        lik = {"kind": "gaussian", "sigma": 0.3}
    elif kind == "bernoulli":
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-2 * f))).astype(float)
        lik = {"kind": "bernoulli", "link": "logit"}
    else:
        y = r.poisson(np.exp(0.6 * f)).astype(float)
        lik = {"kind": "poisson", "link": "log"}
    cf = [{"kind": "sexp", "lscale": 0.9, "magn": 1.1}]
    return x, y, z, cf, lik


def test_oracle_fitc_gaussian_closed_form():
    """FITC Gaussian (R/gp_fit.R:101-123, R/laplace.R:122-152): the log evidence
    must be log N(y | 0, Q + diag(D) + sigma^2 I) with Q = K_xz Kz^-1 K_zx,
    Kz = K_zz + jitter I and D = diag K + jitter - diag Q, from a direct n x n
    Cholesky; the predictive mean and variance (train = FALSE) must be
    Q_tx (Q + L)^-1 y and k_tt - diag(Q_tx (Q + L)^-1 Q_xt), L = diag(D) + sigma^2 I
    (Quinonero-Candela & Rasmussen 2005, eqs. 16 and 24 with the FITC prior)."""
    x, y, z, cf, lik = _fitc_case()
    fit = O.fitc_fit(cf, lik, x, y, z)
    c2 = fit["cfs"]
    n, m, jit = len(y), z.shape[0], 1e-6
    Kz = O.eval_cf(c2, z, z) + jit * np.eye(m)
    Kxz = O.eval_cf(c2, x, z)
    Q = Kxz @ np.linalg.solve(Kz, Kxz.T)
    D = O.eval_cf(c2, x, x, diag_only=True) + jit - np.diag(Q)
    Lam = np.diag(D + 0.09)
    Ky = Q + Lam
    Lc = np.linalg.cholesky(Ky)
    a = np.linalg.solve(Lc, y)
    ref = -0.5 * a @ a - np.sum(np.log(np.diag(Lc))) - 0.5 * n * np.log(2 * np.pi)
    assert abs(fit["log_evidence"] - ref) < 1e-10 * abs(ref)
    xt = np.random.default_rng(2).normal(size=(30, 2))
    Ktz = O.eval_cf(c2, xt, z)
    Qtx = Ktz @ np.linalg.solve(Kz, Kxz.T)
    mu = Qtx @ np.linalg.solve(Ky, y)
    v = O.eval_cf(c2, xt, xt, diag_only=True) - np.sum(Qtx * np.linalg.solve(Ky, Qtx.T).T, axis=1)
    mo, vo = O.fitc_pred(fit, xt, var=True)
    np.testing.assert_allclose(mo, mu, atol=1e-9)
    np.testing.assert_allclose(vo, v, atol=1e-9)
    # full predictive covariance (R/gp_pred.R:258-268): Q_tt + diag(k_tt - diag Q_tt) - Q_tx (Q + L)^-1 Q_xt
    Qtt = Ktz @ np.linalg.solve(Kz, Ktz.T)
    cov_cf = Qtt + np.diag(O.eval_cf(c2, xt, xt, diag_only=True) - np.diag(Qtt)) - Qtx @ np.linalg.solve(Ky, Qtx.T)
    mc, cc = O.fitc_pred_cov(fit, xt, jitter=0.0)
    np.testing.assert_allclose(mc, mu, atol=1e-9)
    np.testing.assert_allclose(cc, cov_cf, atol=1e-9)
    # draws with z = I give the factor of cov + jitter I
    Lf = O.fitc_draw(fit, xt, np.eye(30), jitter=1e-6) - mo[:, None]
    np.testing.assert_allclose(Lf @ Lf.T, cov_cf + 1e-6 * np.eye(30), atol=1e-9)
    assert np.max(np.abs(np.triu(Lf, 1))) < 1e-12


@pytest.mark.parametrize("kind", ["bernoulli", "poisson"])
def test_oracle_fitc_laplace_mode(kind):
    """FITC Laplace: at the mode f = (Q + diag D) grad log p(y | f), i.e. alpha
    equals the likelihood gradient (the Newton fixed point, for any V > 0), and
    the log evidence equals the Laplace formula with the analytic W at the mode
    (Rasmussen & Williams eq. 3.32 with the FITC prior covariance), up to the
    finite-difference curvature of R/lik.R:449-454."""
    x, y, z, cf, lik = _fitc_case(kind=kind)
    fit = O.fitc_fit(cf, lik, x, y, z, maxiter=100, tol=1e-14)
    f = fit["fmean"]
    g = O.loglik_d(lik, f, y)
    np.testing.assert_allclose(fit["alpha"], g, atol=1e-8)
    c2 = fit["cfs"]
    m, jit = z.shape[0], 1e-6
    Kz = O.eval_cf(c2, z, z) + jit * np.eye(m)
    Kxz = O.eval_cf(c2, x, z)
    Q = Kxz @ np.linalg.solve(Kz, Kxz.T)
    Kf = Q + np.diag(O.eval_cf(c2, x, x, diag_only=True) + jit - np.diag(Q))
    np.testing.assert_allclose(Kf @ g, f, atol=1e-8)
    mu = 1 / (1 + np.exp(-f)) if kind == "bernoulli" else np.exp(f)
    W = mu * (1 - mu) if kind == "bernoulli" else mu
    sw = np.sqrt(W)
    B = np.eye(len(y)) + sw[:, None] * Kf * sw[None, :]
    ref = O.loglik(lik, f, y) - 0.5 * f @ g - np.sum(np.log(np.diag(np.linalg.cholesky(B))))
    assert abs(fit["log_evidence"] - ref) < 1e-5 * abs(ref)


def test_oracle_fitc_train_var_closed_form():
    """Training-input variance of FITC (train = TRUE, R/gp_pred.R:270-285): with
    K_t = Q + diag(Dt), Dt = diag K - diag Q (no jitter) and C = Q + diag(D) + sigma^2 I,
    it is diag(K_t - K_t C^-1 K_t) from dense solves."""
    x, y, z, cf, lik = _fitc_case(n=70, m=15)
    fit = O.fitc_fit(cf, lik, x, y, z)
    c2 = fit["cfs"]
    n, m, jit = len(y), z.shape[0], 1e-6
    Kz = O.eval_cf(c2, z, z) + jit * np.eye(m)
    Kxz = O.eval_cf(c2, x, z)
    Q = Kxz @ np.linalg.solve(Kz, Kxz.T)
    kd = O.eval_cf(c2, x, x, diag_only=True)
    Kt = Q + np.diag(kd - np.diag(Q))
    C = Q + np.diag(kd + jit - np.diag(Q) + 0.09)
    ref = np.diag(Kt - Kt @ np.linalg.solve(C, Kt))
    np.testing.assert_allclose(O.fitc_train_var(fit, x), ref, atol=1e-10)


@pytest.mark.parametrize("kind", ["bernoulli", "poisson"])
def test_oracle_fitc_ep_fixed_point(kind):
    """FITC EP (R/ep.R:148-185,332-370): with K_f = Q + diag(D) and the final sites
    (z, V), the posterior mean and variance are K_f (K_f + V)^-1 z and
    diag(K_f - K_f (K_f + V)^-1 K_f), and the log evidence is
    -1/2 log|K_f + V| - 1/2 z^T (K_f + V)^-1 z + sum log Z_i
    + 1/2 sum log(v_cav + V) + 1/2 sum (m_cav - z)^2 / (v_cav + V), all from dense
    solves; at convergence the tilted means match the posterior means (the EP
    fixed point) to the loop tolerance."""
    x, y, z, cf, lik = _fitc_case(kind=kind, n=60, m=10)
    fit = O.fitc_fit(cf, lik, x, y, z, approx="ep", maxiter=200)
    c2 = fit["cfs"]
    m, jit = z.shape[0], 1e-6
    Kz = O.eval_cf(c2, z, z) + jit * np.eye(m)
    Kxz = O.eval_cf(c2, x, z)
    Q = Kxz @ np.linalg.solve(Kz, Kxz.T)
    Kf = Q + np.diag(O.eval_cf(c2, x, x, diag_only=True) + jit - np.diag(Q))
    zz, V = fit["z"], fit["V"]
    Cm = Kf + np.diag(V)
    np.testing.assert_allclose(fit["fmean"], Kf @ np.linalg.solve(Cm, zz), atol=1e-9)
    np.testing.assert_allclose(fit["fvar"], np.diag(Kf - Kf @ np.linalg.solve(Cm, Kf)), atol=1e-9)
    cm, cv = fit["cavity_mean"], fit["cavity_var"]
    zero = np.zeros(len(y))
    mt, _, logZ_site, _, _, _ = O.tilted_moments(lik, cm, 1 / cv, zero, zero, y)  # moments of lik x N(cm, cv)
    sign, ld = np.linalg.slogdet(Cm)
    assert sign > 0
    ref = (-0.5 * ld - 0.5 * zz @ np.linalg.solve(Cm, zz) + np.sum(logZ_site) + 0.5 * np.sum(np.log(cv + V))
           + 0.5 * np.sum((cm - zz) ** 2 / (cv + V)))
    assert abs(fit["log_evidence"] - ref) < 1e-9 * abs(ref)
    np.testing.assert_allclose(mt, fit["fmean"], atol=5e-3)


def test_oracle_loo_gaussian_closed_form():
    """Gaussian likelihood.
    * The LOO latent moments of R/gp_loo.R:120-130 must give the closed form
      s_i^2 = 1/[K_y^-1]_ii and mu_i = y_i - [K_y^-1 y]_i s_i^2 (Rasmussen &
      Williams eqs. 5.10-5.12): loo_var + sigma^2 = s_i^2, loo_mean = mu_i.
    * A 200-node Gauss-Hermite integral over those moments must match
      log N(y_i | mu_i, s_i^2).
    * The 11-node loos (R/quadrature.R's rule through logsumexp) must match the
      same integral computed with numpy's hermegauss(11)."""
    r = np.random.default_rng(8)
    n = 60
    x = r.normal(size=(n, 2))
    y = np.sin(x.sum(1)) + 0.2 * r.normal(size=n)
    cf = [{"kind": "sexp", "lscale": 0.9, "magn": 1.2}]
    lik = {"kind": "gaussian", "sigma": 0.2}
    # jitter 0: with jitter, R's predictive variance at the training inputs uses the
    # jitter-free prior diagonal while the fit's prior carries the jitter, so the
    # identity holds only up to O(jitter) amplified by V/(V - post_var)
    fit = O.gp_fit(cf, lik, x, y, jitter=0.0)
    out = O.gp_loo(fit, lik, x, y, quad_order=11)
    Ky = O.eval_cf(fit["cfs"], x, x) + 0.2 ** 2 * np.eye(n)
    Ki = np.linalg.inv(Ky)
    s2 = 1.0 / np.diag(Ki)
    mu = y - (Ki @ y) * s2
    assert np.max(np.abs(out["loo_var"] + 0.2 ** 2 - s2) / s2) < 1e-8
    assert np.max(np.abs(out["loo_mean"] - mu)) < 1e-8
    ref = -0.5 * np.log(2 * np.pi * s2) - 0.5 * (y - mu) ** 2 / s2

    def gh_loos(q):
        xg, wg = np.polynomial.hermite_e.hermegauss(q)
        wg = wg / wg.sum()
        f = out["loo_mean"][:, None] + np.sqrt(out["loo_var"])[:, None] * xg[None, :]
        ll = -0.5 * np.log(2 * np.pi * 0.04) - 0.5 * (y[:, None] - f) ** 2 / 0.04
        return np.log(np.sum(wg[None, :] * np.exp(ll), axis=1))
    assert np.max(np.abs(gh_loos(200) - ref)) < 1e-9
    assert np.max(np.abs(out["loos"] - gh_loos(11))) < 1e-10


# ---------------------------------------------------------------- beta-binomial and split sites

def test_betabinom_loglik_and_gradient():
    """R/lik.R:332-352 against scipy's beta-binomial pmf; R/lik.R:400-425 against a
    central difference of it."""
    r = _rng(30)
    t = r.integers(1, 30, 40).astype(float)
    y = np.floor(r.uniform(0, 1, 40) * (t + 1))
    f = r.normal(0, 2, 40)
    for link in ("logit", "probit"):
        for phi in (0.05, 1.0, 4.0):
            lik = {"kind": "betabinom", "link": link, "phi": phi}
            mu = O.response(lik, f)
            ref = st.betabinom.logpmf(y, t, mu / phi, (1 - mu) / phi)
            np.testing.assert_allclose(O.loglik(lik, f, y, trials=t, total=False), ref, rtol=1e-11, atol=1e-13)
            h = 1e-6
            fd = (O.loglik(lik, f + h, y, trials=t, total=False) - O.loglik(lik, f - h, y, trials=t, total=False)) / (2 * h)
            np.testing.assert_allclose(O.loglik_d(lik, f, y, trials=t), fd, rtol=1e-5, atol=1e-7)


def _split_case(n=40, nbad=3, seed=31):
    r = _rng(seed)
    x = r.uniform(-2, 2, (n, 1))
    K = O.eval_cf([{"kind": "sexp", "lscale": 0.7, "magn": 1.0}], x, x) + 1e-6 * np.eye(n)
    z = r.normal(size=n)
    V = r.uniform(0.2, 1.0, n)
    bad = r.choice(n, nbad, replace=False)
    V[bad] = -r.uniform(4.0, 8.0, nbad)
    t = np.full(n, 10.0)
    y = np.floor(r.uniform(0, 11, n))
    return K, z, V, y, t, np.flatnonzero(V > 0), np.sort(bad)


def _dense_split(K, z, V, good, bad):
    """The split branch's Gaussian pieces by dense solves: m, P = the bad latents'
    predictive given the good pseudo-data, f2 = their posterior mean under the bad
    pseudo-data, 1/2 log|I + V_bad^-1 P|."""
    M11 = K[np.ix_(good, good)] + np.diag(V[good])
    K21 = K[np.ix_(bad, good)]
    m = K21 @ np.linalg.solve(M11, z[good])
    P = K[np.ix_(bad, bad)] - K21 @ np.linalg.solve(M11, K21.T)
    f2 = m + P @ np.linalg.solve(P + np.diag(V[bad]), z[bad] - m)
    sgn, ld = np.linalg.slogdet(np.eye(len(bad)) + P / V[bad][:, None])
    assert sgn > 0
    return m, P, f2, 0.5 * ld


def test_split_laplace_step_matches_dense_formulas():
    """R/laplace.R:40-117 on sites with three negative pseudo-variances: the evidence
    p(y_good) p(y_bad | y_good) and alpha from dense solves and determinants (no
    Cholesky chain, no LU), and alpha = (K + V)^-1 z."""
    K, z, V, y, t, good, bad = _split_case()
    lik = {"kind": "betabinom", "link": "logit", "phi": 0.5}
    fit = O.laplace_iter(lik, K, y, None, pobs=(z, V), trials=t)
    assert "C_lu" in fit
    K11 = K[np.ix_(good, good)]
    f1 = K11 @ np.linalg.solve(K11 + np.diag(V[good]), z[good])
    le1 = (O.loglik(lik, f1, y[good], trials=t[good]) - 0.5 * f1 @ np.linalg.solve(K11, f1)
           - 0.5 * np.linalg.slogdet(np.eye(len(good)) + K11 / V[good][None, :])[1])
    m, P, f2, half_logdet = _dense_split(K, z, V, good, bad)
    le2 = (O.loglik(lik, f2, y[bad], trials=t[bad]) - 0.5 * (f2 - m) @ np.linalg.solve(P, f2 - m) + half_logdet)
    assert fit["log_evidence"] == pytest.approx(le1 + le2, rel=1e-9)
    alpha = np.linalg.solve(K + np.diag(V), z)
    np.testing.assert_allclose(fit["alpha"], alpha, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(fit["fmean"], K @ alpha, rtol=1e-9, atol=1e-12)


def test_split_ep_sweep_matches_dense_formulas():
    """R/ep.R:50-138 with the same split: the bad sites' evidence terms with
    (m - z)^T (P + V_bad)^-1 (m - z) and 1/2 log|I + V_bad^-1 P| from dense
    algebra, and fvar = diag(V (K + V)^-1 K)."""
    K, z, V, y, t, good, bad = _split_case(seed=32)
    r = _rng(33)
    n = len(z)
    logZs = r.normal(-1, 0.3, n)
    cm = r.normal(size=n)
    cv = r.uniform(0.5, 2.0, n)  # below |V| at the bad sites, so (c + V) / V > 0 as in an EP state
    fit = O.ep_iter({"kind": "betabinom", "link": "logit", "phi": 0.5}, K, y, None, None, None, None,
                    pobs=(z, V, logZs, cm, cv, True), trials=t)
    assert "C_lu" in fit
    K11 = K[np.ix_(good, good)]
    M11 = K11 + np.diag(V[good])
    zg = z[good]
    le1 = (-0.5 * np.linalg.slogdet(M11)[1] - 0.5 * zg @ np.linalg.solve(M11, zg) + np.sum(logZs[good])
           + 0.5 * np.sum(np.log(cv[good] + V[good])) + 0.5 * np.sum((cm[good] - zg) ** 2 / (cv[good] + V[good])))
    m, P, _, half_logdet = _dense_split(K, z, V, good, bad)
    d = m - z[bad]
    le2 = (-0.5 * d @ np.linalg.solve(P + np.diag(V[bad]), d) + half_logdet + np.sum(logZs[bad])
           + 0.5 * np.sum(np.log((cv[bad] + V[bad]) / V[bad])) + 0.5 * np.sum((cm[bad] - z[bad]) ** 2 / (cv[bad] + V[bad])))
    assert fit["log_evidence"] == pytest.approx(le1 + le2, rel=1e-9)
    np.testing.assert_allclose(fit["fvar"], V * np.diag(np.linalg.solve(K + np.diag(V), K)), rtol=1e-9, atol=1e-12)


def _betabinom_outliers(seed=3, n=100, phi=0.2):
    r = _rng(seed)
    x = np.sort(r.uniform(-3, 3, n))
    f = 3.0 + 0.5 * np.sin(x)
    mu = 1 / (1 + np.exp(-f))
    y = r.binomial(10, r.beta(mu / phi, (1 - mu) / phi)).astype(float)
    y[[10, 50, 90]] = 0.0
    return x[:, None], y, np.full(n, 10.0)


def test_split_laplace_mode_is_the_fixed_point():
    """With beta-binomial outliers the converged Newton iterate splits its sites
    (some pseudo-variances are negative at the mode), and it is still the mode:
    f = K grad log p(y | f), which holds for any V."""
    x, y, t = _betabinom_outliers()
    lik = {"kind": "betabinom", "link": "logit", "phi": 0.2}
    fit = O.gp_fit([{"kind": "sexp", "lscale": 1.0, "magn": 1.0}], lik, x, y, trials=t, maxiter=100, tol=1e-10)
    assert "C_lu" in fit and np.sum(fit["V"] < 0) >= 1
    f = fit["fmean"]
    np.testing.assert_allclose(f, fit["K"] @ O.loglik_d(lik, f, y, trials=t), atol=1e-8)
