"""The benchmark workloads at their full sizes, against the oracle.

* n = 8192, d = 8, Matern-5/2 ARD, sigma = 0.05 (energy + gradient, prediction at
  4096 points): the plain FP64 device path against the oracle's FP64 log evidence
  and gradient (the CPU restatement on host LAPACK) and its predictive mean and
  variance at every test point, relative 1e-8.
* n = 4096, d = 4, Bernoulli-logit Laplace, tolerance 1e-8: both forms take the
  oracle's Newton step count; mode and log evidence to 1e-8 (FP64) and 1e-6
  (mixed).
* 2048 batched regressions (n log-uniform in [64, 2048], d in 1..8), six
  refinement sweeps: every problem converges; every energy against the plain
  FP64 device fit of the same problem (1e-6 of the log density's summands);
  16 sampled problems, including the largest n and the largest n with d = 8,
  against the oracle's energy and gradient (gradient 1e-3, from the FP32 B^-1).
"""

import warnings

import numpy as np
import pytest

import oracle as O
from tests.test_gpu_mixed import logz_tol

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def c2():
    from gplite_amd import synth
    return synth.c2()


def test_c2_fp64_energy_gradient_and_prediction_match_oracle(gpl, c2):
    G = gpl
    from oracle.cpu_baseline import energy_grad_matern52_ard
    w = c2
    gp = G.gp_init(G.cf_matern52(lscale=list(w["lscale"]), magn=w["magn"]), G.lik_gaussian(sigma=w["sigma"]))
    gf = G.gp_fit(gp, w["x"], w["y"], want_grad=True)
    g = G.gp_energy_grad(gf)
    pr = G.gp_pred(gf, w["xt"], var=True)
    lz_o, g_o, _ = energy_grad_matern52_ard(w["x"], w["y"], w["lscale"], w["magn"], w["sigma"])
    print(f"C2 fp64: logZ {gf.fit.log_evidence!r} vs {lz_o!r}; grad rel {rel(g, g_o):.2e}", flush=True)
    assert gf.fit.log_evidence == pytest.approx(lz_o, rel=1e-8)
    assert rel(g, g_o) < 1e-8
    fo = O.gp_fit([c.spec() for c in gf.cfs], gf.lik.spec(), w["x"], w["y"])
    assert fo["log_evidence"] == pytest.approx(lz_o, rel=1e-10)
    mo, vo = O.gp_pred(fo, w["xt"], var=True)
    print(f"C2 fp64 pred: mean rel {rel(pr['mean'], mo):.2e}, var rel {rel(pr['var'], vo):.2e}", flush=True)
    assert rel(pr["mean"], mo) < 1e-8
    assert rel(pr["var"], vo) < 1e-8
    assert pr["var"].shape == (4096,)


@pytest.fixture(scope="module")
def c3_oracle():
    from gplite_amd import synth
    import gplite_amd as G
    w = synth.c3()
    gp = G.gp_init(G.cf_sexp(lscale=w["lscale"], magn=w["magn"]), G.lik_bernoulli("logit"),
                   approx=G.approx_laplace(maxiter=30, tol=1e-8))
    with warnings.catch_warnings():
        warnings.simplefilter("error")  # the oracle must converge within the cap
        fo = O.gp_fit([c.spec() for c in gp.cfs], gp.lik.spec(), w["x"], w["y"], tol=1e-8)
    return w, gp, fo


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-8), ("mixed", 1e-6)])
def test_c3_laplace_matches_oracle(gpl, c3_oracle, prec, tol):
    G = gpl
    w, gp, fo = c3_oracle
    gf = G.gp_fit(gp, w["x"], w["y"], precision=prec)
    print(f"C3 {prec}: steps {gf.fit.iters} vs {fo['iters']}; logZ {gf.fit.log_evidence!r} vs "
          f"{fo['log_evidence']!r}; mode rel {rel(gf.fit.fmean, fo['fmean']):.2e}", flush=True)
    assert gf.fit.iters == fo["iters"]
    assert gf.fit.log_evidence == pytest.approx(fo["log_evidence"], rel=tol)
    assert rel(gf.fit.fmean, fo["fmean"]) < tol
    assert rel(gf.fit.alpha, fo["alpha"]) < tol * 10
    if prec == "mixed":
        assert not gf.fit.precision_info["refit_fp64"]


def test_c4_full_batch(gpl):
    G = gpl
    from gplite_amd import _lib as L
    from gplite_amd import synth
    from gplite_amd.batch import GPBatch
    probs = synth.c4()
    gps = [G.gp_init(G.cf_sexp(lscale=p["lscale"], magn=p["magn"]), G.lik_gaussian(sigma=p["sigma"])) for p in probs]
    B = GPBatch(gps, [p["x"] for p in probs], [p["y"] for p in probs], refine_iters=6)
    e, g = B.energy_grad(include_prior=False)
    st, rr = B.status()
    assert np.all(st == L.GPL_OK) and np.all(np.isfinite(e))
    print(f"C4: {len(probs)} problems, largest last refinement correction {rr.max():.2e}", flush=True)
    # every problem against the plain FP64 device fit
    worst = 0.0
    for i, p in enumerate(probs):
        gi = G.gp_fit(B.gps[i], p["x"], p["y"])
        eo = -gi.fit.log_evidence
        tol = logz_tol(len(p["y"]), p["sigma"], eo)
        worst = max(worst, abs(e[i] - eo) / tol)
        assert abs(e[i] - eo) <= tol, (i, e[i], eo)
    print(f"C4: worst |E - E_fp64| / tolerance {worst:.3f}", flush=True)
    # a sample against the oracle, with the largest problems
    ns = np.array([p["x"].shape[0] for p in probs])
    ds = np.array([p["x"].shape[1] for p in probs])
    pick = {int(np.argmax(ns)), int(np.flatnonzero(ds == 8)[np.argmax(ns[ds == 8])])}
    rr2 = np.random.default_rng(4)
    while len(pick) < 16:
        pick.add(int(rr2.integers(0, len(probs))))
    for i in sorted(pick):
        p = probs[i]
        fo = O.gp_fit([c.spec() for c in B.gps[i].cfs], B.gps[i].lik.spec(), p["x"], p["y"])
        eo = O.gp_energy(fo, include_prior=False)
        assert abs(e[i] - eo) <= logz_tol(len(p["y"]), p["sigma"], eo), (i, e[i], eo)
        go = O.gp_energy_grad(fo, include_prior=False)
        assert np.max(np.abs(g[i] - go)) <= 1e-3 * np.max(np.abs(go)), i
