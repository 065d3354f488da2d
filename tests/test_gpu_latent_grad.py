"""The Laplace fit with the exact curvature and the gradient of Laplace and EP fits.

The oracle restates Rasmussen & Williams (2006) Algorithm 5.1 for the Laplace
gradient (explicit term plus the mode's dependence through d3 log p) and the
EP fixed-point gradient (eq. 5.27); both are pinned by central differences in
tests/test_oracle.py.  Here the device (through the C ABI) is held to the oracle
at 1e-8 relative on the log evidence, the mode, the step count and the
gradient, for the logit, probit and log links, Bernoulli and binomial; and at
n = 1024 the device gradient is held to central differences of the device's own
log evidence (a size-independent check of the whole chain: B^-1, the rank-two
weight term, the fused gradient contraction).
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


LIKS = [("bernoulli", "logit"), ("bernoulli", "probit"), ("poisson", "log"), ("binomial", "logit"),
        ("binomial", "probit")]


def _lik(G, kind, link):
    return {"bernoulli": G.lik_bernoulli, "binomial": G.lik_binomial, "poisson": G.lik_poisson}[kind](link)


def _data(kind, n, d=2, seed=0):
    r = np.random.default_rng(seed)
    x = r.uniform(-2, 2, (n, d))
    f = np.sin(x[:, 0]) + 0.5 * x[:, -1]
    tr = None
    if kind == "bernoulli":
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-2 * f))).astype(float)
    elif kind == "poisson":
        y = r.poisson(np.exp(f)).astype(float)
    else:
        tr = r.integers(1, 8, n).astype(float)
        y = r.binomial(tr.astype(int), 1 / (1 + np.exp(-f))).astype(float)
    return x, y, tr


def _oracle(gp, x, y, tr, approx, **kw):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return O.gp_fit([c.spec() for c in gp.cfs], gp.lik.spec(), x, y, approx=approx, trials=tr, **kw)


@pytest.mark.parametrize("kind,link", LIKS)
def test_laplace_with_exact_curvature_matches_oracle(gpl, kind, link):
    G = gpl
    x, y, tr = _data(kind, 200, seed=1)
    gp = G.gp_init(G.cf_sexp(lscale=[0.8, 1.2], magn=1.1), _lik(G, kind, link),
                   approx=G.approx_laplace(analytic_w=True))
    gf = G.gp_fit(gp, x, y, trials=tr)
    fo = _oracle(gp, x, y, tr, "laplace", analytic_w=True)
    assert gf.fit.iters == fo["iters"]
    assert gf.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=RTOL)
    assert rel(gf.fit.fmean, fo["fmean"]) < RTOL
    # and it differs from the reference's forward-difference curvature at the 1e-6 level
    fd = _oracle(gp, x, y, tr, "laplace")
    assert 0 < abs(fd["log_evidence"] - fo["log_evidence"]) < 1e-4 * abs(fo["log_evidence"])


@pytest.mark.parametrize("approx", ["laplace", "ep"])
@pytest.mark.parametrize("kind,link", LIKS)
def test_latent_gradient_matches_oracle(gpl, kind, link, approx):
    G = gpl
    x, y, tr = _data(kind, 150, d=3, seed=2)
    cfs = [G.cf_matern32(lscale=[0.9, 1.4, 0.7], magn=1.2), G.cf_const(magn=0.5)]
    ap = G.approx_ep() if approx == "ep" else G.approx_laplace(analytic_w=True)
    gp = G.gp_init(cfs, _lik(G, kind, link), approx=ap)
    gf = G.gp_fit(gp, x, y, trials=tr)
    g = G.gp_energy_grad(gf, include_prior=False)
    fo = _oracle(gp, x, y, tr, approx, analytic_w=(approx == "laplace"))
    go = O.gp_energy_grad(fo, include_prior=False)
    print(f"{approx} {kind}/{link}: grad rel {rel(g, go):.2e}", flush=True)
    assert gf.fit.iters == fo["iters"]
    assert g.shape == go.shape
    assert rel(g, go) < RTOL
    # the prior term is added on the host as for the Gaussian likelihood
    np.testing.assert_allclose(G.gp_energy_grad(gf) - g, -O.lpdf_prior_grad(fo["cfs"], fo["lik"]), rtol=1e-12)


@pytest.mark.parametrize("approx", ["laplace", "ep"])
def test_latent_gradient_matches_central_differences_at_n1024(gpl, approx):
    """Size-independent: the device gradient against central differences of the
    device's own log evidence (Newton to 1e-11 for Laplace; EP at the reference's
    tolerance, where the fixed-point gradient holds to ~1e-4)."""
    G = gpl
    x, y, tr = _data("bernoulli", 1024, d=4, seed=3)

    def fit(ls, magn):
        ap = G.approx_ep() if approx == "ep" else G.approx_laplace(maxiter=100, tol=1e-11, analytic_w=True)
        gp = G.gp_init(G.cf_sexp(lscale=list(ls), magn=magn), G.lik_bernoulli(), approx=ap)
        return G.gp_fit(gp, x, y)

    th = np.array([0.9, 1.3, 0.8, 1.1, 1.4])
    gf = fit(th[:4], th[4])
    g = G.gp_energy_grad(gf, include_prior=False)
    h = 1e-5
    fd = []
    for j in range(5):
        tp = th.copy(); tp[j] *= np.exp(h)
        tm = th.copy(); tm[j] *= np.exp(-h)
        fd.append(-(fit(tp[:4], tp[4]).fit.log_evidence - fit(tm[:4], tm[4]).fit.log_evidence) / (2 * h))
    fd = np.array(fd)
    print(f"{approx} n=1024: grad vs central differences rel {rel(g, fd):.2e}", flush=True)
    assert rel(g, fd) < (1e-6 if approx == "laplace" else 1e-3)


def test_latent_gradient_refuses_what_it_does_not_cover(gpl):
    G = gpl
    from gplite_amd import _lib as L
    x, y, tr = _data("bernoulli", 100)
    gp = G.gp_init(G.cf_sexp(), G.lik_bernoulli())  # forward-difference curvature
    gf = G.gp_fit(gp, x, y)
    with pytest.raises(L.GplError) as ei:
        G.gp_energy_grad(gf)
    assert ei.value.code == L.GPL_ERR_UNSUPPORTED
    gm = G.gp_fit(G.gp_init(G.cf_sexp(), G.lik_bernoulli(), approx=G.approx_laplace(analytic_w=True)), x, y,
                  precision="mixed")
    with pytest.raises(L.GplError) as ei:
        G.gp_energy_grad(gm)
    assert ei.value.code == L.GPL_ERR_UNSUPPORTED
    t = np.full(100, 10.0)
    with pytest.raises(L.GplError) as ei:
        G.gp_fit(G.gp_init(G.cf_sexp(), G.lik_betabinom(), approx=G.approx_laplace(analytic_w=True)), x,
                 np.minimum(y * 10, 10), trials=t)
    assert ei.value.code == L.GPL_ERR_UNSUPPORTED
    # beta-binomial EP: the gradient in phi is not provided unless phi is fixed
    ge = G.gp_fit(G.gp_init(G.cf_sexp(), G.lik_betabinom(), approx=G.approx_ep()), x, np.minimum(y * 10, 10),
                  trials=t)
    with pytest.raises(NotImplementedError):
        G.gp_energy_grad(ge)


def test_gradient_optimizer_on_a_classification_problem(gpl):
    """gp_optim's L-BFGS-B path on a Laplace (exact curvature) fit reaches a
    stationary point of the energy."""
    G = gpl
    x, y, tr = _data("bernoulli", 300, d=2, seed=4)
    gp = G.gp_init(G.cf_sexp(lscale=[1.0, 1.0], magn=1.0), G.lik_bernoulli(),
                   approx=G.approx_laplace(maxiter=100, tol=1e-10, analytic_w=True))
    go = G.gp_optim(gp, x, y, method="L-BFGS-B", restarts=0, maxiter=200, tol=1e-10, verbose=False)
    g = G.gp_energy_grad(go)
    assert np.max(np.abs(g)) / len(y) < 1e-4
