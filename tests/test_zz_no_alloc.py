"""gpl_pred_noalloc: prediction whose scratch is a fixed static device array.

Called through the C ABI and compared with gpl_pred on the same model and test
points, bit for bit (the same kernels; test points are processed in chunks and
every output column is computed independently of the others), on two calls in
a row, the second with fewer test points.  Covers the FP64 model with the
blocked triangular solve, the FP64 model with L^-1 (several chunks), the mixed
model (FP32 L^-1), and the size limit.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pred_both(G, L, gf, xt, var=True):
    _, _, xtp, _ = G.api._program(gf.cfs, xt)
    m = xtp.shape[0]
    mean_a, var_a = np.empty(m), np.empty(m)
    L.check(L.lib().gpl_pred(gf.fit.handle, L.dptr(xtp), m, int(var), L.dptr(mean_a), L.dptr(var_a) if var else None))
    mean_b, var_b = np.full(m, np.nan), np.full(m, np.nan)
    L.check(L.lib().gpl_pred_noalloc(gf.fit.handle, L.dptr(xtp), m, int(var), L.dptr(mean_b),
                                     L.dptr(var_b) if var else None))
    return (mean_a, var_a), (mean_b, var_b)


def _check(G, L, gf, xts):
    for xt in xts:  # two calls in a row, the second smaller
        (ma, va), (mb, vb) = _pred_both(G, L, gf, xt)
        print(f"n={gf.fit.n} m={xt.shape[0]}: max |mean diff| {np.max(np.abs(ma - mb)):.1e}", flush=True)
        assert np.array_equal(ma, mb)
        assert np.array_equal(va, vb)
        (ma, _), (mb, _) = _pred_both(G, L, gf, xt, var=False)
        assert np.array_equal(ma, mb)


def test_noalloc_matches_pred_fp64_blocked_solve(gpl):
    """FP64 Laplace model (no L^-1: the blocked triangular solve)."""
    G = gpl
    from gplite_amd import _lib as L
    r = np.random.default_rng(1)
    x = r.normal(size=(700, 2))
    y = (r.uniform(size=700) < 1 / (1 + np.exp(-np.sin(2 * x[:, 0])))).astype(float)
    gf = G.gp_fit(G.gp_init(G.cf_sexp(lscale=0.7, magn=1.5), G.lik_bernoulli("logit")), x, y)
    _check(G, L, gf, [r.normal(size=(1000, 2)), r.normal(size=(37, 2))])


def test_noalloc_matches_pred_in_chunks(gpl):
    """FP64 Gaussian model at n = 20000 with L^-1: 16 N bytes per test point leave
    room for 192 points per chunk in the 64 MiB array, so 1000 points take six."""
    G = gpl
    from gplite_amd import _lib as L
    r = np.random.default_rng(2)
    x = r.uniform(0, 1, (20000, 3))
    y = np.sin(3 * x.sum(1)) + 0.1 * r.standard_normal(20000)
    gf = G.gp_fit(G.gp_init(G.cf_matern52(lscale=[0.3, 0.4, 0.5]), G.lik_gaussian(sigma=0.1)), x, y,
                  want_grad=True)
    _check(G, L, gf, [r.uniform(0, 1, (1000, 3)), r.uniform(0, 1, (150, 3))])


def test_noalloc_matches_pred_mixed(gpl):
    G = gpl
    from gplite_amd import _lib as L
    r = np.random.default_rng(3)
    x = r.uniform(0, 1, (3000, 4))
    y = np.cos(2 * x.sum(1)) + 0.1 * r.standard_normal(3000)
    gf = G.gp_fit(G.gp_init(G.cf_matern32(lscale=0.6), G.lik_gaussian(sigma=0.1)), x, y, precision="mixed")
    assert gf.fit.precision_info["precision"] == "mixed"
    _check(G, L, gf, [r.uniform(0, 1, (2500, 4)), r.uniform(0, 1, (10, 4))])


def test_noalloc_rejects_models_beyond_its_scratch(gpl):
    """n = 65500 pads to 65536, beyond the 65408 that 64 test points can use in
    64 MiB: a clear GPL_ERR_ARG, and gpl_pred still serves the model."""
    G = gpl
    from gplite_amd import _lib as L
    r = np.random.default_rng(4)
    n = 65500
    x = r.uniform(0, 1, (n, 1))
    y = np.sin(6 * x[:, 0]) + 0.1 * r.standard_normal(n)
    gf = G.gp_fit(G.gp_init(G.cf_sexp(lscale=0.2), G.lik_gaussian(sigma=0.1)), x, y)
    xt = r.uniform(0, 1, (5, 1))
    mean, var = np.empty(5), np.empty(5)
    rc = L.lib().gpl_pred_noalloc(gf.fit.handle, L.dptr(xt), 5, 1, L.dptr(mean), L.dptr(var))
    assert rc == L.GPL_ERR_ARG
    assert "too large" in L.lib().gpl_last_error().decode()
    assert np.all(np.isfinite(G.gp_pred(gf, xt, var=True)["var"]))
