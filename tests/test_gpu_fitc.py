"""FITC (method_fitc: R/gp_fit.R:101-123, R/laplace.R:122-152,300-322,
R/gp_pred.R:234-290) on the device against the oracle's restatement with the
same inducing inputs: log evidence, posterior mean, alpha, pseudo-data and the
FITC diagonal at relative 1e-8, the Newton step count, predictive mean and
variance at 1e-8 and joint draws at 1e-7 (a Cholesky of the predictive
covariance, whose smallest eigenvalues are near the jitter), for Gaussian, Bernoulli-logit, binomial (trials) and Poisson
(offset) likelihoods over ragged n and m.  Inducing inputs placed by the
package's k-means are checked the same way (the oracle reads them back).  A
factor failure gives log_evidence = -Inf as in the reference."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _data(kind, n, m, seed):
    r = np.random.default_rng(seed)
    x = r.normal(size=(n, 2))
    z = r.normal(size=(m, 2)) * 1.3
    f = np.sin(x.sum(1))
    kw = {}
    if kind == "gaussian":
        y = f + 0.3 * r.normal(size=n)
    elif kind == "logit":
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-2 * f))).astype(float)
    elif kind == "binomial":
        t = r.integers(1, 6, size=n).astype(float)
        y = r.binomial(t.astype(int), 1 / (1 + np.exp(-f))).astype(float)
        kw["trials"] = t
    else:
        off = 0.3 * r.normal(size=n)
        y = r.poisson(np.exp(0.5 * f + off)).astype(float)
        kw["offset"] = off
    return x, y, z, kw


def _lik(G, kind):
    return {"gaussian": G.lik_gaussian(sigma=0.3), "logit": G.lik_bernoulli("logit"),
            "binomial": G.lik_binomial("logit"), "poisson": G.lik_poisson()}[kind]


@pytest.mark.parametrize("kind,n,m", [("gaussian", 1, 1), ("gaussian", 150, 20), ("gaussian", 333, 70),
                                      ("logit", 200, 30), ("binomial", 150, 25), ("poisson", 130, 65),
                                      # long reductions: the K_zx S^-1 K_xz product is split over n (4 and 2 parts)
                                      ("gaussian", 4096, 64), ("logit", 2048, 40)])
def test_fitc_matches_oracle(gpl, kind, n, m):
    G = gpl
    x, y, z, kw = _data(kind, n, m, 100 + n + m)
    lk = _lik(G, kind)
    gp = G.gp_init(G.cf_matern52(lscale=[0.9, 1.2], magn=1.1), lk, method=G.method_fitc(inducing=z),
                   approx=G.approx_laplace(maxiter=50, tol=1e-8))
    gf = G.gp_fit(gp, x, y, **kw)
    fo = O.fitc_fit([c.spec() for c in gf.cfs], lk.spec(), x, y, z, maxiter=50, tol=1e-8, **kw)
    assert gf.fit.iters == fo["iters"]
    assert abs(gf.fit.log_evidence - fo["log_evidence"]) <= 1e-8 * abs(fo["log_evidence"])
    for nm, ref in (("fmean", fo["fmean"]), ("alpha", fo["alpha"]), ("z", fo["z"]), ("V", fo["V"]),
                    ("diag", fo["D"])):
        assert _rel(getattr(gf.fit, nm), ref) < 1e-8, nm
    xt = np.random.default_rng(5).normal(size=(77, 2))
    pr = G.gp_pred(gf, xt, var=True)
    mo, vo = O.fitc_pred(fo, xt, var=True)
    assert _rel(pr["mean"], mo) < 1e-8
    assert _rel(pr["var"], vo) < 1e-8
    # joint draws (R/gp_draw.R:126-140) on the same standard normal matrix
    d = G.gp_draw(gf, xt, draws=3, transform=False, seed=11)
    zz = np.random.default_rng(11).standard_normal((77, 3))
    ref = O.fitc_draw(fo, xt, zz)
    assert _rel(d, ref) < 1e-7


def test_fitc_kmeans_inducing(gpl):
    G = gpl
    x, y, _, _ = _data("gaussian", 400, 1, 9)
    lk = _lik(G, "gaussian")
    gp = G.gp_init(G.cf_sexp(lscale=0.8), lk, method=G.method_fitc(num_inducing=24, bin_along=0, bin_count=4))
    gf = G.gp_fit(gp, x, y)
    z = gf.method.inducing
    assert z.shape == (24, 2) and gp.method.inducing is None  # the input model is not modified
    fo = O.fitc_fit([c.spec() for c in gf.cfs], lk.spec(), x, y, z)
    assert abs(gf.fit.log_evidence - fo["log_evidence"]) <= 1e-8 * abs(fo["log_evidence"])
    pr = G.gp_pred(gf, x[:50], var=True)
    mo, vo = O.fitc_pred(fo, x[:50], var=True)
    assert _rel(pr["mean"], mo) < 1e-8 and _rel(pr["var"], vo) < 1e-8


def test_fitc_failed_factor_gives_minus_inf(gpl, capsys):
    G = gpl
    x, y, z, _ = _data("gaussian", 60, 8, 3)
    gp = G.gp_init(G.cf_sexp(magn=0.5), G.lik_gaussian(sigma=0.3), method=G.method_fitc(inducing=z))
    gf = G.gp_fit(gp, x, y, jitter=-1.0)  # Kz = K_zz - I is indefinite
    assert gf.fit.log_evidence == -np.inf
    assert "not positive" in capsys.readouterr().out


@pytest.mark.parametrize("kind,n,m", [("logit", 120, 15), ("poisson", 140, 30), ("binomial", 100, 20)])
def test_fitc_ep_matches_oracle(gpl, kind, n, m):
    """FITC with approx_ep (R/ep.R:148-185,332-370): log evidence, posterior mean and
    variance, cavity moments and the step count against the oracle at 1e-8."""
    G = gpl
    x, y, z, kw = _data(kind, n, m, 300 + n)
    lk = _lik(G, kind)
    gp = G.gp_init(G.cf_sexp(lscale=[0.9, 1.2], magn=1.1), lk, method=G.method_fitc(inducing=z),
                   approx=G.approx_ep(damping=0.8))
    gf = G.gp_fit(gp, x, y, **kw)
    fo = O.fitc_fit([c.spec() for c in gf.cfs], lk.spec(), x, y, z, approx="ep", damping=0.8, **kw)
    assert gf.fit.iters == fo["iters"]
    assert abs(gf.fit.log_evidence - fo["log_evidence"]) <= 1e-8 * abs(fo["log_evidence"])
    for nm, ref in (("fmean", fo["fmean"]), ("fvar", fo["fvar"]), ("cavity_mean", fo["cavity_mean"]),
                    ("cavity_var", fo["cavity_var"]), ("alpha", fo["alpha"])):
        assert _rel(getattr(gf.fit, nm), ref) < 1e-8, nm
    xt = np.random.default_rng(6).normal(size=(40, 2))
    pr = G.gp_pred(gf, xt, var=True)
    mo, vo = O.fitc_pred(fo, xt, var=True)
    assert _rel(pr["mean"], mo) < 1e-8 and _rel(pr["var"], vo) < 1e-8


@pytest.mark.parametrize("kind,approx", [("gaussian", "laplace"), ("logit", "laplace"), ("poisson", "laplace"),
                                         ("logit", "ep"), ("poisson", "ep")])
def test_fitc_loo_matches_oracle(gpl, kind, approx):
    """gp_loo for method_fitc (R/gp_loo.R:73-134): the Laplace LOO moments use the
    training-input variance with the FITC diagonal (train = TRUE), EP the cavities."""
    G = gpl
    x, y, z, kw = _data(kind, 130, 25, 500)
    lk = _lik(G, kind)
    ap = G.approx_ep() if approx == "ep" else G.approx_laplace(maxiter=50, tol=1e-8)
    gp = G.gp_init(G.cf_matern32(lscale=1.0, magn=1.0), lk, method=G.method_fitc(inducing=z), approx=ap)
    gf = G.gp_fit(gp, x, y, **kw)
    okw = dict(approx="ep") if approx == "ep" else dict(maxiter=50, tol=1e-8)
    fo = O.fitc_fit([c.spec() for c in gf.cfs], lk.spec(), x, y, z, **okw, **kw)
    ld = G.gp_loo(gf, x, y, quad_order=11, **kw)
    lo = O.gp_loo(fo, lk.spec(), x, y, quad_order=11, **kw)
    assert _rel(ld["loos"], lo["loos"]) < 1e-8
    assert ld["loo"] == pytest.approx(lo["loo"], rel=1e-8)
