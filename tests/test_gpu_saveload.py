"""gp_save / gp_load (R/saveload.R:32-41) of fitted models: the loaded model is
rebuilt on the device from the saved gp$fit fields (gpl_model_restore, no refit)
and must behave like the model that was saved.

* fit fields: vectors bit-identical, C_chol to 1e-14 relative (it is rescaled by
  1/sqrt(V) on load and by sqrt(V) on export), log evidence identical;
* FP64 fits (Gaussian, Bernoulli-logit Laplace, Poisson EP): predictive mean and
  variance against the oracle at 1e-8 relative, gp_loo, gp_draw and the energy
  gradient against the saved model at 1e-8;
* a mixed-precision fit loads as FP64: mean against the oracle at 1e-6, variance
  at 1e-3 (the saved factor is the FP32 one);
* a saved V <= 0 is an argument error, a non-positive C_chol diagonal NOT_PD."""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _case(G, case, n):
    r = np.random.default_rng(40 + n)
    x = r.normal(size=(n, 2))
    f = np.sin(x.sum(1))
    if case in ("gaussian", "gaussian_mixed"):
        lk, approx, y = G.lik_gaussian(sigma=0.3), G.approx_laplace(), f + 0.3 * r.normal(size=n)
        kw = dict(maxiter=50, tol=1e-4)
    elif case == "logit_laplace":
        lk, approx = G.lik_bernoulli("logit"), G.approx_laplace(maxiter=50, tol=1e-8)
        y = (r.uniform(size=n) < 1 / (1 + np.exp(-2 * f))).astype(float)
        kw = dict(maxiter=50, tol=1e-8)
    else:
        lk, approx = G.lik_poisson(), G.approx_ep()
        y = r.poisson(np.exp(0.5 * f)).astype(float)
        kw = dict(approx="ep")
    gp = G.gp_init(G.cf_matern52(lscale=[0.9, 1.2], magn=1.1), lk, approx=approx)
    prec = "mixed" if case == "gaussian_mixed" else None
    gf = G.gp_fit(gp, x, y, precision=prec)
    fo = O.gp_fit([c.spec() for c in gf.cfs], lk.spec(), x, y, **kw)
    return gf, fo, x, y


@pytest.mark.parametrize("case,n", [("gaussian", 1), ("gaussian", 200), ("logit_laplace", 150),
                                    ("poisson_ep", 130), ("gaussian_mixed", 300)])
def test_save_load_round_trip(gpl, tmp_path, case, n):
    G = gpl
    gf, fo, x, y = _case(G, case, n)
    path = str(tmp_path / "fit.model")
    G.gp_save(gf, path)
    g2 = G.gp_load(path)
    assert G.is_fitted(g2)
    assert g2.fit.log_evidence == gf.fit.log_evidence
    assert G.gp_energy(g2) == G.gp_energy(gf)
    for nm in ("fmean", "alpha", "z", "V", "fvar"):
        np.testing.assert_array_equal(getattr(g2.fit, nm), getattr(gf.fit, nm))
    c1, c2 = gf.fit.C_chol, g2.fit.C_chol
    assert _rel(c2, c1) <= 1e-14
    xt = np.random.default_rng(7).normal(size=(70, 2))
    pr = G.gp_pred(g2, xt, var=True)
    mo, vo = O.gp_pred(fo, xt, var=True)
    if case == "gaussian_mixed":
        assert _rel(pr["mean"], mo) < 1e-6
        assert _rel(pr["var"], vo) < 1e-3
        return
    assert _rel(pr["mean"], mo) < 1e-8
    assert _rel(pr["var"], vo) < 1e-8
    d1 = G.gp_draw(gf, xt, draws=3, transform=False, seed=2)
    d2 = G.gp_draw(g2, xt, draws=3, transform=False, seed=2)
    assert _rel(d2, d1) < 1e-8
    if n > 1:
        l1 = G.gp_loo(gf, x, y)
        l2 = G.gp_loo(g2, x, y)
        assert _rel(l2["loos"], l1["loos"]) < 1e-8
    if case == "gaussian":
        assert _rel(G.gp_energy_grad(g2), G.gp_energy_grad(gf)) < 1e-8


def _resave(path, out, **change):
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    arrays.update(change)
    with open(out, "wb") as fh:
        np.savez(fh, **arrays)


def test_load_rejects_bad_fit_fields(gpl, tmp_path):
    G = gpl
    from gplite_amd import _lib as L
    gf, _, _, _ = _case(G, "gaussian", 100)
    path = str(tmp_path / "fit.model")
    G.gp_save(gf, path)
    with np.load(path, allow_pickle=False) as z:
        V, C = z["V"].copy(), z["C_chol"].copy()
    V[5] = 0.0
    _resave(path, str(tmp_path / "badv.model"), V=V)
    with pytest.raises(L.GplError) as ei:
        G.gp_load(str(tmp_path / "badv.model"))
    assert ei.value.code == L.GPL_ERR_ARG
    C[70, 70] = -abs(C[70, 70])
    _resave(path, str(tmp_path / "badc.model"), C_chol=C)
    with pytest.raises(L.GplError) as ei:
        G.gp_load(str(tmp_path / "badc.model"))
    assert ei.value.code == L.GPL_ERR_NOT_PD
    assert ei.value.minor == 71


def test_loaded_latent_models_and_the_gradient(gpl, tmp_path):
    """An EP fit's gradient needs only the fit fields: the loaded model gives the
    saved model's gradient (1e-8).  The Laplace gradient (exact curvature) also
    reads the targets, which a saved fit does not hold: the loaded model refuses
    it, and a refit on the same data gives it again."""
    G = gpl
    from gplite_amd import _lib as L
    r = np.random.default_rng(8)
    x = r.normal(size=(150, 2))
    f = np.sin(x.sum(1))
    y = (r.uniform(size=150) < 1 / (1 + np.exp(-2 * f))).astype(float)
    ge = G.gp_fit(G.gp_init(G.cf_sexp(lscale=[0.9, 1.1], magn=1.3), G.lik_bernoulli(), approx=G.approx_ep()), x, y)
    G.gp_save(ge, str(tmp_path / "ep.model"))
    ge2 = G.gp_load(str(tmp_path / "ep.model"))
    assert _rel(G.gp_energy_grad(ge2), G.gp_energy_grad(ge)) < 1e-8
    gl = G.gp_fit(G.gp_init(G.cf_sexp(lscale=[0.9, 1.1], magn=1.3), G.lik_bernoulli(),
                            approx=G.approx_laplace(analytic_w=True)), x, y)
    G.gp_save(gl, str(tmp_path / "la.model"))
    gl2 = G.gp_load(str(tmp_path / "la.model"))
    with pytest.raises(L.GplError) as ei:
        G.gp_energy_grad(gl2)
    assert ei.value.code == L.GPL_ERR_UNSUPPORTED
    gl3 = G.gp_fit(gl2, x, y)
    assert _rel(G.gp_energy_grad(gl3), G.gp_energy_grad(gl)) < 1e-12
