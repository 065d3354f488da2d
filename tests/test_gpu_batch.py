"""Batched form (many independent Gaussian GPs of uneven n in one pass) against
the FP64 oracle per problem.  Stated tolerances as for the mixed-precision
form: the log evidence within 1e-6 relative to the size of its summands
(max(|logZ|, n(|log sigma| + log(2 pi)/2)); the FP32 trailing updates
perturb log|B| by an absolute amount that does not shrink with |logZ|), the
gradient within 1e-3 relative."""

import math

import numpy as np
import pytest

import oracle as O
from tests.test_gpu_core import GRAD_CASES


def logz_tol(n, sigma, logz, rtol=1e-6):
    return rtol * max(abs(logz), n * (abs(math.log(sigma)) + 0.5 * math.log(2 * math.pi)))

pytestmark = pytest.mark.gpu


def _problems(G, nprob=24, seed=3):
    r = np.random.default_rng(seed)
    gps, xs, ys = [], [], []
    for p in range(nprob):
        n = int(r.integers(1, 400)) if p else 1
        d = int(r.integers(1, 5))
        x = r.uniform(-2, 2, (n, d))
        y = np.sin(x.sum(1)) + 0.1 * r.standard_normal(n)
        gps.append(G.gp_init(G.cf_sexp(lscale=float(r.uniform(0.3, 1.5)), magn=1.0), G.lik_gaussian(sigma=0.1)))
        xs.append(x)
        ys.append(y)
    return gps, xs, ys


def test_batch_matches_oracle_per_problem(gpl):
    G = gpl
    from gplite_amd.batch import GPBatch
    gps, xs, ys = _problems(G)
    B = GPBatch(gps, xs, ys, refine_iters=6)
    e, g = B.energy_grad()
    for p in range(len(gps)):
        fo = O.gp_fit([c.spec() for c in B.gps[p].cfs], B.gps[p].lik.spec(), xs[p], ys[p])
        eo = O.gp_energy(fo)
        assert abs(e[p] - eo) <= logz_tol(len(ys[p]), 0.1, eo), (p, e[p], eo)
        go = O.gp_energy_grad(fo)
        assert np.max(np.abs(g[p] - go)) <= 1e-3 * np.max(np.abs(go)), (p, g[p], go)
    # graph replay is deterministic
    e2, g2 = B.energy_grad()
    assert np.array_equal(e, e2) and np.array_equal(g, g2)
    assert B.info()["buckets"] >= 2


def test_batch_params_and_tuning(gpl):
    G = gpl
    from gplite_amd.batch import GPBatch
    gps, xs, ys = _problems(G, nprob=10, seed=9)
    B = GPBatch(gps, xs, ys)
    table = B.tune()
    assert len(table) == B.info()["buckets"] and all(t["tile"] in (64, 128) for t in table)
    B.load_tuning(table)
    p0 = [np.array(list(G.get_param(g).values())) + 0.2 for g in B.gps]
    B.set_params(p0)
    e, g = B.energy_grad()
    for p in range(len(gps)):
        fo = O.gp_fit([c.spec() for c in B.gps[p].cfs], B.gps[p].lik.spec(), xs[p], ys[p])
        eo = O.gp_energy(fo)
        assert abs(e[p] - eo) <= logz_tol(len(ys[p]), B.gps[p].lik.sigma, eo), (p, e[p], eo)


def test_batch_reports_failed_factor_per_problem(gpl):
    """A problem whose B has a negative pivot (negative jitter on a tiny signal)
    gets energy +Inf while the others keep their results; without allow_not_pd the
    call raises naming that problem."""
    G = gpl
    from gplite_amd._lib import GplError
    from gplite_amd.batch import GPBatch
    r = np.random.default_rng(2)
    x = r.uniform(-2, 2, (100, 2))
    y = np.sin(x.sum(1))
    good = G.gp_init(G.cf_sexp(lscale=1.0, magn=1.0), G.lik_gaussian(sigma=1.0))
    bad = G.gp_init(G.cf_sexp(lscale=1.0, magn=1e-3), G.lik_gaussian(sigma=0.1))
    B = GPBatch([good, bad, good], [x, x, x], [y, y, y], jitter=-0.5)
    e, g = B.energy_grad(allow_not_pd=True)
    assert np.isfinite(e[0]) and np.isfinite(e[2]) and e[1] == np.inf
    assert np.all(np.isnan(g[1])) and np.all(np.isfinite(g[0]))
    assert e[0] == e[2]
    with pytest.raises(GplError) as ei:
        B.energy_grad()
    assert "problem 1" in str(ei.value)


def test_batch_set_params_rebuilds_periodic_features(gpl):
    """A periodic term's features [sin(2 pi x / period), cos(...)] depend on the
    period, so a parameter update rebuilds them from the resident inputs: after
    set_params with new periods and length-scales every problem's energy and
    gradient match the oracle at the new parameters, and equal a batch created
    directly at them."""
    G = gpl
    from gplite_amd.batch import GPBatch
    r = np.random.default_rng(77)
    gps, xs, ys = [], [], []
    for p, n in enumerate((60, 200, 333)):
        x = r.normal(size=(n, 2))
        ys.append(np.sin(2.0 * x[:, 0]) + 0.3 * x[:, 1] + 0.1 * r.normal(size=n))
        xs.append(x)
        cfs = [G.cf_periodic(vars=[0], period=2.0 + 0.3 * p, cf_base=G.cf_sexp(lscale=0.8)), G.cf_lin(vars=[1])]
        gps.append(G.gp_init(cfs, G.lik_gaussian(sigma=0.2)))
    B = GPBatch(gps, xs, ys, refine_iters=6)
    e0, _ = B.energy_grad()
    newp = [np.array(list(G.get_param(g).values())) + 0.15 * r.normal(size=len(G.get_param(g))) for g in B.gps]
    B.set_params(newp)
    e, g = B.energy_grad()
    assert not np.allclose(e, e0)
    B2 = GPBatch([G.set_param(gp, pp) for gp, pp in zip(gps, newp)], xs, ys, refine_iters=6)
    e2, g2 = B2.energy_grad()
    np.testing.assert_array_equal(e, e2)
    np.testing.assert_array_equal(g, g2)
    for p in range(3):
        fo = O.gp_fit([c.spec() for c in B.gps[p].cfs], B.gps[p].lik.spec(), xs[p], ys[p])
        eo = O.gp_energy(fo)
        assert abs(e[p] - eo) <= logz_tol(len(ys[p]), B.gps[p].lik.sigma, eo), (p, e[p], eo)
        go = O.gp_energy_grad(fo)
        assert np.max(np.abs(g[p] - go)) <= 1e-3 * np.max(np.abs(go)), p


@pytest.mark.parametrize("ci", range(len(GRAD_CASES)))
def test_batch_covariance_structures(gpl, ci):
    """Every covariance structure of the FP64 gradient tests (sums, products,
    nested products, periodic, nn, lin) in one batch of three problems of
    uneven n with their own hyperparameters: energy and gradient per problem
    against the oracle, as in test_batch_matches_oracle_per_problem."""
    G = gpl
    from gplite_amd.batch import GPBatch
    r = np.random.default_rng(300 + ci)
    gps, xs, ys = [], [], []
    for p, n in enumerate((37, 150, 301)):
        x = r.normal(size=(n, 2))
        ys.append(np.sin(x[:, 0]) * np.cos(0.5 * x[:, 1]) + 0.1 * r.normal(size=n))
        xs.append(x)
        gp = G.gp_init(GRAD_CASES[ci](G), G.lik_gaussian(sigma=0.1 + 0.05 * p))
        pp = np.array(list(G.get_param(gp).values()))
        gps.append(G.set_param(gp, pp + 0.1 * r.normal(size=pp.size)))
    B = GPBatch(gps, xs, ys, refine_iters=6)
    e, g = B.energy_grad()
    for p in range(3):
        fo = O.gp_fit([c.spec() for c in B.gps[p].cfs], B.gps[p].lik.spec(), xs[p], ys[p])
        eo = O.gp_energy(fo)
        assert abs(e[p] - eo) <= logz_tol(len(ys[p]), B.gps[p].lik.sigma, eo), (ci, p, e[p], eo)
        go = O.gp_energy_grad(fo)
        assert np.max(np.abs(g[p] - go)) <= 1e-3 * np.max(np.abs(go)), (ci, p)
