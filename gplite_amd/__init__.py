"""gplite_amd: MI355X-native dense Gaussian-process core with gplite's API.

The heavy lifting (covariance builds, Cholesky, solves, gradient, prediction,
Laplace and EP loops) runs in libgplite_amd.so (hand-written gfx950 HIP
kernels behind a C ABI, include/gplite_amd.h).  This package mirrors the R
front end.
"""

from . import _lib
from .api import (  # noqa: F401
    GP, approx_ep, approx_laplace, cf_const, cf_lin, cf_matern32, cf_matern52, cf_nn, cf_periodic, cf_prod,
    cf_sexp, eval_cf, get_param, gp_compare, gp_draw, gp_energy, gp_energy_grad, gp_fit, gp_init, gp_load, gp_loo, gp_optim,
    gp_pred, gp_save, is_fitted, lik_bernoulli, lik_betabinom, lik_binomial, lik_gaussian, lik_poisson, lpdf_prior, method_fitc,
    method_full, prior_fixed, prior_half_t, prior_lognormal, prior_logunif, set_param,
)

__version__ = "0.1.0"


def require_device():
    """Load the HIP library and open GPU 0; raises if either is unavailable."""
    _lib.lib()
    _lib.ctx()
    return True
