"""Serving form: results alone equal results in a concurrent mix (continuous
batching with several workers and coalesced predictions), and footprint
admission with eviction keeps working under a small budget."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(G, k=6, seed=1):
    r = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(r.integers(50, 700))
        x = r.uniform(0, 1, (n, 3))
        f = np.sin(3 * x.sum(1))
        if i % 2:
            gp = G.gp_init(G.cf_matern32(lscale=0.4), G.lik_poisson())
            y = r.poisson(np.exp(f)).astype(float)
        else:
            gp = G.gp_init(G.cf_matern32(lscale=0.4), G.lik_gaussian(sigma=0.1))
            y = f + 0.1 * r.standard_normal(n)
        xts = [r.uniform(0, 1, (int(r.integers(1, 300)), 3)) for _ in range(3)]
        out.append((gp, x, y, xts))
    return out


def test_alone_equals_mix(gpl):
    G = gpl
    from gplite_amd.serve import Server
    reqs = _requests(G)
    # alone: one worker, every request waited for before the next
    s1 = Server(workers=1, mem_gb=8)
    alone = []
    for gp, x, y, xts in reqs:
        mid, tk = s1.submit_fit(gp, x, y)
        lz = s1.wait(tk)["log_evidence"]
        preds = []
        for xt in xts:
            preds.append(s1.wait(s1.submit_pred(mid, xt)))
        alone.append((lz, preds))
    s1.close()
    # mix: four workers, everything submitted at once
    s4 = Server(workers=4, mem_gb=8)
    fit_t, pred_t = [], []
    for gp, x, y, xts in reqs:
        mid, tk = s4.submit_fit(gp, x, y)
        fit_t.append(tk)
        pred_t.append([s4.submit_pred(mid, xt) for xt in xts])
    coalesced = 0
    for i, (lz, preds) in enumerate(alone):
        assert s4.wait(fit_t[i])["log_evidence"] == lz
        for tk, pa in zip(pred_t[i], preds):
            pm = s4.wait(tk)
            coalesced = max(coalesced, pm["coalesced"])
            assert np.array_equal(pm["mean"], pa["mean"]) and np.array_equal(pm["var"], pa["var"])
    st = s4.stats()
    s4.close()
    assert st["fits"] == len(reqs) and st["preds"] == 3 * len(reqs)
    assert coalesced >= 1


def test_small_budget_evicts_idle_models(gpl):
    G = gpl
    from gplite_amd.serve import Server
    reqs = _requests(G, k=5, seed=7)
    # ~6.4 MB: the two largest fitted models (about 4.5 and 3.5 MB) cannot both stay resident
    s = Server(workers=2, mem_gb=0.006)
    lz = []
    for gp, x, y, xts in reqs:
        mid, tk = s.submit_fit(gp, x, y)
        lz.append(s.wait(tk)["log_evidence"])
    st = s.stats()
    assert st["evictions"] >= 1 and st["bytes_in_use"] <= 0.006 * 2 ** 30
    s.close()
    assert all(np.isfinite(lz))


def _c5_request(G, r, n, poisson, d=6):
    """A fit request of the serving workload's shape: x ~ U[0,1]^6, cf_matern32 with
    lscale ~ U[0.2, 0.8]; Gaussian (sigma 0.1) or Poisson counts with a Laplace fit."""
    x = r.uniform(0, 1, (n, d))
    f = np.sin(2.0 * x.sum(1)) + 0.5 * np.cos(3.0 * x[:, 0])
    cf = G.cf_matern32(lscale=float(r.uniform(0.2, 0.8)), magn=1.0)
    if poisson:
        return G.gp_init(cf, G.lik_poisson()), x, r.poisson(np.exp(f)).astype(float)
    return G.gp_init(cf, G.lik_gaussian(sigma=0.1)), x, f + 0.1 * r.standard_normal(n)


def _rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("precision", ["fp64", "auto"])
def test_served_results_match_direct_fits_and_the_oracle(gpl, precision):
    """Every served fit and prediction of a concurrent C5-shaped mix (n from 128 to
    2048, both likelihoods) equals the same request made directly on the library's
    own context (bitwise: the same kernels on other streams): gp_fit, or for the
    Gaussian fits that "auto" runs in the mixed form, the batched fit the server
    gives them (gp_fit_batch); and the oracle at the stated tolerances: FP64 1e-8,
    mixed log evidence 1e-6 of its summands, predictive mean 1e-6, variance 1e-3."""
    import oracle as O
    from gplite_amd.serve import Server
    from tests.test_gpu_mixed import logz_tol
    G = gpl
    r = np.random.default_rng(55)
    reqs = [_c5_request(G, r, n, poisson=i % 2 == 1) for i, n in enumerate([128, 2048, 700, 1500, 333, 1024])]
    xts = [r.uniform(0, 1, (1000, 6)) for _ in reqs]
    s = Server(workers=4, mem_gb=16)
    tickets = []
    for (gp, x, y), xt in zip(reqs, xts):
        mid, tk = s.submit_fit(gp, x, y, precision=precision)
        tickets.append((tk, s.submit_pred(mid, xt), s.submit_pred(mid, xt[:257])))
    for i, ((gp, x, y), xt) in enumerate(zip(reqs, xts)):
        fr = s.wait(tickets[i][0])
        pr, pr2 = s.wait(tickets[i][1]), s.wait(tickets[i][2])
        gauss = isinstance(gp.lik, G.lik_gaussian)
        prec = "mixed" if (precision == "auto" and gauss) else "fp64"
        if prec == "mixed":
            from gplite_amd.batch import gp_fit_batch
            gd = gp_fit_batch([gp], [x], [y], refine_iters=8)[0]
        else:
            gd = G.gp_fit(gp, x, y, precision=prec)
        assert fr["log_evidence"] == gd.fit.log_evidence, i
        pd = G.gp_pred(gd, xt, var=True)
        assert np.array_equal(pr["mean"], pd["mean"]) and np.array_equal(pr["var"], pd["var"]), i
        assert np.array_equal(pr2["mean"], pr["mean"][:257]) and np.array_equal(pr2["var"], pr["var"][:257]), i
        fo = O.gp_fit([c.spec() for c in gd.cfs], gd.lik.spec(), x, y)
        mo, vo = O.gp_pred(fo, xt, var=True)
        if prec == "fp64":
            assert fr["log_evidence"] == pytest.approx(fo["log_evidence"], rel=1e-8), i
            assert _rel(pr["mean"], mo) < 1e-8 and _rel(pr["var"], vo) < 1e-8, i
        else:
            assert abs(fr["log_evidence"] - fo["log_evidence"]) <= logz_tol(len(y), 0.1, fo["log_evidence"]), i
            assert _rel(pr["mean"], mo) < 1e-6 and _rel(pr["var"], vo) < 1e-3, i
    s.close()


def test_c5_largest_requests_alone_equal_mix(gpl):
    """The serving workload's largest shapes: a Poisson-Laplace fit at n = 16384 and a
    Gaussian one at n = 12000 among small requests, with m = 1000 predictions.  Each
    result is the same alone (one worker, one request at a time) and in the
    concurrent mix (four workers, everything submitted at once)."""
    from gplite_amd.serve import Server
    G = gpl
    r = np.random.default_rng(77)
    reqs = [_c5_request(G, r, n, poisson=p) for n, p in
            [(16384, True), (200, False), (12000, False), (150, True), (900, True)]]
    xts = [r.uniform(0, 1, (1000, 6)) for _ in reqs]
    outs = []
    for workers, serial in ((1, True), (4, False)):
        s = Server(workers=workers, mem_gb=64)
        res, pend = [], []
        for (gp, x, y), xt in zip(reqs, xts):
            mid, tk = s.submit_fit(gp, x, y, precision="auto")
            if serial:
                lz = s.wait(tk)["log_evidence"]
                p = s.wait(s.submit_pred(mid, xt))
                res.append((lz, p["mean"], p["var"]))
            else:
                pend.append((tk, s.submit_pred(mid, xt)))
        for tk, tp in pend:
            lz = s.wait(tk)["log_evidence"]
            p = s.wait(tp)
            res.append((lz, p["mean"], p["var"]))
        s.close()
        outs.append(res)
    for (a, b) in zip(*outs):
        assert np.isfinite(a[0]) and a[0] == b[0]
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_evicted_model_is_refitted_for_a_prediction(gpl):
    """With room for one model, fitting a second evicts the first; a prediction on the
    first then refits it from its kept inputs and returns the result it gave before."""
    from gplite_amd.serve import Server
    G = gpl
    r = np.random.default_rng(8)
    (g1, x1, y1), (g2, x2, y2) = _c5_request(G, r, 600, False), _c5_request(G, r, 640, True)
    xt = r.uniform(0, 1, (100, 6))
    s = Server(workers=2, mem_gb=0.0093)  # ~10 MB: the two models (about 4.6 and 7.9 MB) do not fit together
    m1, t1 = s.submit_fit(g1, x1, y1)
    lz1 = s.wait(t1)["log_evidence"]
    p1 = s.wait(s.submit_pred(m1, xt))
    m2, t2 = s.submit_fit(g2, x2, y2)
    s.wait(t2)
    st = s.stats()
    assert st["evictions"] >= 1 and st["resident_models"] == 1
    p1b = s.wait(s.submit_pred(m1, xt))
    assert np.array_equal(p1["mean"], p1b["mean"]) and np.array_equal(p1["var"], p1b["var"])
    assert s.stats()["fits"] == 3 and np.isfinite(lz1)
    s.close()


def test_release_is_final(gpl):
    """A released model id is unknown afterwards: a second release and a prediction fail
    cleanly, and the server keeps serving."""
    from gplite_amd import _lib as L
    from gplite_amd.serve import Server
    G = gpl
    r = np.random.default_rng(9)
    gp, x, y = _c5_request(G, r, 300, False)
    s = Server(workers=2, mem_gb=1)
    mid, tk = s.submit_fit(gp, x, y)
    pt = s.submit_pred(mid, x[:10])
    s.release(mid)  # waits for the fit and the prediction
    assert np.all(np.isfinite(s.wait(pt)["mean"]))
    s._keep[("model", mid)] = gp
    with pytest.raises(L.GplError):
        s.release(mid)
    with pytest.raises(L.GplError):
        s.submit_pred(mid, x[:10])
    mid2, tk2 = s.submit_fit(gp, x, y)
    assert np.isfinite(s.wait(tk2)["log_evidence"])
    s.close()


@pytest.mark.parametrize("kind", ["laplace", "mixed", "ep", "laplace_exact"])
def test_trimmed_model_predicts_and_differentiates_as_before(gpl, kind):
    """gpl_model_trim (what the server does to every resident model) keeps what
    prediction, draws and the gradient read: the same results, bit for bit, with
    fewer device bytes."""
    from gplite_amd import _lib as L
    G = gpl
    r = np.random.default_rng(10)
    if kind == "mixed":
        gp, x, y = _c5_request(G, r, 900, False)
        gf = G.gp_fit(gp, x, y, precision="mixed")
    else:
        gp, x, y = _c5_request(G, r, 900, True)
        if kind == "ep":
            gp = G.gp_init(gp.cfs, gp.lik, approx=G.approx_ep())
        if kind == "laplace_exact":
            gp = G.gp_init(gp.cfs, gp.lik, approx=G.approx_laplace(analytic_w=True))
        gf = G.gp_fit(gp, x, y)
    xt = r.uniform(0, 1, (200, 6))
    # the gradient first: it may form L^-1, which prediction then uses in place of the
    # blocked triangular solve (equal to rounding, not bit for bit)
    g0 = G.gp_energy_grad(gf) if kind in ("mixed", "ep", "laplace_exact") else None
    before = G.gp_pred(gf, xt, var=True)
    d0 = G.gp_draw(gf, xt[:20], draws=3, transform=False, seed=4) if kind != "ep" else None
    b0 = L.lib().gpl_model_bytes(gf.fit.handle)
    L.check(L.lib().gpl_model_trim(gf.fit.handle))
    assert L.lib().gpl_model_bytes(gf.fit.handle) < b0
    after = G.gp_pred(gf, xt, var=True)
    assert np.array_equal(before["mean"], after["mean"]) and np.array_equal(before["var"], after["var"])
    if d0 is not None:
        assert np.array_equal(d0, G.gp_draw(gf, xt[:20], draws=3, transform=False, seed=4))
    if g0 is not None:
        assert np.array_equal(g0, G.gp_energy_grad(gf))
