"""Failure reporting of a single-problem fit.  In the reference a failed
Cholesky of B raises R's "the leading minor of order k is not positive" and
method_full has no tryCatch (R/gp_fit.R:91-99,150-161), so the error reaches
the caller; here gpl_fit returns GPL_ERR_NOT_PD with k (gpl_last_error_minor).

The case has a known failing minor: with cf_lin (K = x x^T), sigma = 1 and
jitter -2, B = I + K - 2 I = x x^T - I.  With x_1 = 3, x_2 = 2 the first pivot is
8 and the second 3 - 36/8 = -1.5, so LAPACK's dpotrf stops at order 2; the
device must name the same minor in both precisions, and the model handle given
for a refit must stay usable."""

import numpy as np
import pytest
from scipy.linalg import lapack

pytestmark = pytest.mark.gpu


def _case():
    r = np.random.default_rng(3)
    n = 100
    x = r.uniform(-0.1, 0.1, n)
    x[0], x[1] = 3.0, 2.0
    y = r.standard_normal(n)
    B = np.outer(x, x) - np.eye(n)
    info = lapack.dpotrf(B, lower=1)[1]
    assert info == 2
    return x, y, info


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_gaussian_not_pd_names_the_minor(gpl, prec):
    G = gpl
    from gplite_amd import _lib as L
    x, y, info = _case()
    gp = G.gp_init(G.cf_lin(magn=1.0), G.lik_gaussian(sigma=1.0))
    with pytest.raises(L.GplError) as ei:
        G.gp_fit(gp, x, y, jitter=-2.0, precision=prec)
    assert ei.value.code == L.GPL_ERR_NOT_PD
    assert ei.value.minor == info
    assert "leading minor of order 2" in str(ei.value)
    # the failure leaves nothing behind: a good fit follows on the same context
    gf = G.gp_fit(gp, x, y, precision=prec)
    assert np.isfinite(gf.fit.log_evidence)


def test_laplace_not_pd_raises(gpl):
    """Laplace factors B = I + S K S with the same indefinite K (S from the first
    sites: B = -11.5 I + x x^T / 4 at f = 0, so the first pivot is already
    negative); the failure is reported, not returned as a number."""
    G = gpl
    from gplite_amd import _lib as L
    x, _, _ = _case()
    yb = (np.arange(x.size) % 2).astype(float)
    gp = G.gp_init(G.cf_lin(magn=1.0), G.lik_bernoulli("logit"), approx=G.approx_laplace())
    with pytest.raises(L.GplError) as ei:
        G.gp_fit(gp, x, yb, jitter=-50.0)
    assert ei.value.code == L.GPL_ERR_NOT_PD
    assert ei.value.minor == 1


def test_ep_negative_cavity_gives_minus_inf(gpl):
    """EP starts from zero site precisions, so its first posterior variances are
    diag(K) < 0 here and the tilted moments are NaN: the reference returns
    log_evidence = -Inf (R/ep.R:322-328), and so does the device fit."""
    G = gpl
    x, _, _ = _case()
    yb = (np.arange(x.size) % 2).astype(float)
    gp = G.gp_init(G.cf_lin(magn=1.0), G.lik_bernoulli("logit"), approx=G.approx_ep())
    gf = G.gp_fit(gp, x, yb, jitter=-50.0)
    assert gf.fit.log_evidence == -np.inf
