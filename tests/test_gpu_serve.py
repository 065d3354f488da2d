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
