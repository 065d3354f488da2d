"""Mixed-precision form (FP32 MFMA storage/updates, FP64 diagonal blocks,
log-determinant and refinement) against the FP64 oracle.

Stated tolerances: relative 1e-6 on the log evidence and the predictive mean
(the precision-form target); the gradient comes from an FP32 B^-1 and is held
to 1e-3 relative; the predictive variance to 1e-4 relative.
"""

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("n,d,sigma", [(64, 1, 0.1), (333, 2, 0.1), (1500, 8, 0.05), (2048, 8, 0.05)])
def test_mixed_gaussian_parity(gpl, n, d, sigma):
    G = gpl
    r = np.random.default_rng(n + d)
    x = r.uniform(0, 1, (n, d))
    y = np.sin(3 * x.sum(1)) + sigma * r.normal(size=n)
    ls = r.uniform(0.2, 1.0, d)
    gp = G.gp_init(G.cf_matern52(lscale=ls, magn=1.0), G.lik_gaussian(sigma=sigma))
    gm = G.gp_fit(gp, x, y, precision="mixed")
    fo = O.gp_fit([c.spec() for c in gm.cfs], gm.lik.spec(), x, y)
    assert gm.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=1e-6)
    assert rel(gm.fit.alpha, fo["alpha"]) < 1e-6
    xt = r.uniform(0, 1, (257, d))
    pr = G.gp_pred(gm, xt, var=True)
    mo, vo = O.gp_pred(fo, xt, var=True)
    assert rel(pr["mean"], mo) < 1e-6
    g = G.gp_energy_grad(gm)
    go = O.gp_energy_grad(fo)
    assert rel(g, go) < 1e-3
    pr = G.gp_pred(gm, xt, var=True)
    assert rel(pr["var"], vo) < 1e-4


def test_mixed_matches_plain_device_path(gpl):
    """The two device forms agree on the same inputs (FP64 plain as reference)."""
    G = gpl
    r = np.random.default_rng(5)
    n = 1024
    x = r.normal(size=(n, 3)); y = np.cos(x.sum(1)) + 0.1 * r.normal(size=n)
    gp = G.gp_init(G.cf_sexp(lscale=[0.6, 0.8, 1.0], magn=1.3), G.lik_gaussian(sigma=0.1))
    g64 = G.gp_fit(gp, x, y, want_grad=True)
    g32 = G.gp_fit(gp, x, y, precision="mixed")
    assert g32.fit.log_evidence == pytest.approx(g64.fit.log_evidence, rel=1e-6)
    assert rel(G.gp_energy_grad(g32), G.gp_energy_grad(g64)) < 1e-3
