"""gp_save / gp_load (R/saveload.R:32-41) of an unfitted model: the stored
structure (covariance functions with ARD scales, products, periodic bases,
learned normalisation, likelihoods, priors, approximations) round-trips exactly,
and the file is an .npz archive that numpy reads without pickle."""

import numpy as np
import pytest

import gplite_amd as G


def _models():
    cf1 = G.cf_sexp(vars=[0, 1], lscale=[0.7, 1.3], magn=0.9, prior_magn=G.prior_fixed())
    cf1.normalize = True
    cf1.learn_scales(np.random.default_rng(1).normal(size=(20, 3)))
    yield G.gp_init([cf1, G.cf_periodic(vars=[2], period=2.5, cf_base=G.cf_matern32(lscale=0.4)),
                     G.cf_const(magn=0.3)], G.lik_gaussian(sigma=0.2, prior_sigma=G.prior_lognormal(0.1, 2.0)))
    yield G.gp_init(G.cf_nn(sigma0=0.5) * G.cf_lin(magn=1.5, prior_magn=G.prior_fixed()),
                    G.lik_bernoulli("probit"), approx=G.approx_ep(damping=0.8, quad_order=9))
    yield G.gp_init(G.cf_matern52(), G.lik_poisson(), approx=G.approx_laplace(maxiter=40, tol=1e-7))


@pytest.mark.parametrize("k", range(3))
def test_unfitted_round_trip(tmp_path, k):
    gp = list(_models())[k]
    f = tmp_path / "gp.model"
    G.gp_save(gp, str(f))
    assert f.exists()  # the name is kept as given (no suffix appended)
    with np.load(str(f), allow_pickle=False) as z:
        assert "header" in z.files
    g2 = G.gp_load(str(f))
    assert not G.is_fitted(g2)
    assert type(g2.lik) is type(gp.lik) and type(g2.approx) is type(gp.approx)
    assert [c.spec() for c in g2.cfs].__repr__() == [c.spec() for c in gp.cfs].__repr__()
    assert g2.lik.spec() == gp.lik.spec()
    assert vars(g2.approx) == vars(gp.approx)
    np.testing.assert_array_equal(list(G.get_param(g2).values()), list(G.get_param(gp).values()))
    assert list(G.get_param(g2).keys()) == list(G.get_param(gp).keys())
    assert G.lpdf_prior(g2) == G.lpdf_prior(gp)


def test_load_rejects_foreign_classes(tmp_path):
    import json
    f = tmp_path / "bad.model"
    head = {"format": "gplite_amd-gp-1", "cfs": [{"__class__": "Popen", "state": {}}],
            "lik": None, "method": None, "approx": None, "precision": "fp64", "fitted": False}
    with open(f, "wb") as fh:
        np.savez(fh, header=np.array(json.dumps(head)))
    with pytest.raises(ValueError, match="unknown object class"):
        G.gp_load(str(f))
