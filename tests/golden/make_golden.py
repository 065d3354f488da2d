"""Writes tests/golden/c1.npz: the CPU-runnable workload (n=200, cf_sexp(0.3,1),
lik_gaussian(0.1), pred at 300 points in [-4,4]) with the FP64 oracle's
outputs, after checking them against scipy's multivariate-normal density and
the closed-form predictive equations.  Also a small Bernoulli-logit Laplace
and EP case (n=60).  R itself is absent here, so these are oracle vectors
pinned by closed forms, not outputs of the R package."""

import os
import sys

import numpy as np
import scipy.stats as st

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from gplite_amd import synth  # noqa: E402


def c1_case():
    w = synth.c1()
    cfs = [{"kind": "sexp", "lscale": w["lscale"], "magn": w["magn"]}]
    lik = {"kind": "gaussian", "sigma": w["sigma"]}
    fit = O.gp_fit(cfs, lik, w["x"], w["y"])
    K = O.eval_cf(cfs, w["x"], w["x"])
    n = len(w["y"])
    ref = st.multivariate_normal(np.zeros(n), K + (w["sigma"] ** 2 + O.JITTER) * np.eye(n)).logpdf(w["y"])
    assert abs(fit["log_evidence"] / ref - 1) < 1e-12
    mean, var = O.gp_pred(fit, w["xt"], var=True)
    Kt = O.eval_cf(cfs, w["xt"], w["x"])
    C = K + (w["sigma"] ** 2 + O.JITTER) * np.eye(n)
    assert np.allclose(mean, Kt @ np.linalg.solve(C, w["y"]), rtol=1e-10, atol=1e-12)
    grad = O.gp_energy_grad(fit)
    return dict(c1_x=w["x"], c1_y=w["y"], c1_xt=w["xt"], c1_logz=fit["log_evidence"],
                c1_energy=O.gp_energy(fit), c1_alpha=fit["alpha"], c1_fmean=fit["fmean"],
                c1_mean=mean, c1_var=var, c1_grad=grad)


def class_case():
    r = np.random.default_rng(6060)
    n = 60
    x = r.standard_normal((n, 2))
    f = np.sin(2 * x[:, 0]) + 0.5 * x[:, 1]
    y = (r.uniform(size=n) < 1 / (1 + np.exp(-f))).astype(float)
    cfs = [{"kind": "matern32", "lscale": 0.9, "magn": 1.5}]
    lik = {"kind": "bernoulli", "link": "logit"}
    la = O.gp_fit(cfs, lik, x, y, maxiter=50, tol=1e-8)
    ep = O.gp_fit(cfs, lik, x, y, approx="ep")
    return dict(cl_x=x, cl_y=y, cl_la_logz=la["log_evidence"], cl_la_fmean=la["fmean"],
                cl_la_iters=la["iters"], cl_ep_logz=ep["log_evidence"], cl_ep_fmean=ep["fmean"],
                cl_ep_fvar=ep["fvar"], cl_ep_iters=ep["iters"])


def build():
    out = {}
    out.update(c1_case())
    out.update(class_case())
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1.npz")
    np.savez(path, **build())
    print("wrote", path)
