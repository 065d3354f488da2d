"""The bounded waits between the workgroups of the single-launch Cholesky panel
and of the single-launch triangular solve.

With the poll limit forced to 0 (GPL_PANEL_WAIT_POLLS, read when a context is
created), the workgroups below a diagonal block give up their first wait. Every
workgroup then leaves the kernel, and the fit reports a device error instead of a
result: Gaussian fits in FP64 and in the mixed form, a Laplace fit (Cholesky with
the forward solve carried, then the single-launch backward solve) and a FITC fit
(single-launch solves in both directions on the inducing-point factors). The
same fits in a fresh process with the default limit succeed. Each case runs in
a child process of its own, so the override never reaches the other tests'
contexts.
"""

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CODE = """
import numpy as np
import gplite_amd as G
from gplite_amd import _lib as L
r = np.random.default_rng(0)
x = r.uniform(size=(1000, 2))
y = np.sin(6 * x.sum(1)) + 0.1 * r.standard_normal(1000)
cases = [("fp64", G.gp_init(G.cf_sexp(lscale=[0.3, 0.3]), G.lik_gaussian(sigma=0.1)), y, {"precision": "fp64"}),
         ("mixed", G.gp_init(G.cf_sexp(lscale=[0.3, 0.3]), G.lik_gaussian(sigma=0.1)), y, {"precision": "mixed"}),
         ("laplace", G.gp_init(G.cf_sexp(lscale=[0.3, 0.3]), G.lik_bernoulli("logit")), (y > 0).astype(float),
          {"precision": "fp64"}),
         ("fitc", G.gp_init(G.cf_sexp(lscale=[0.3, 0.3]), G.lik_gaussian(sigma=0.1),
                            method=G.method_fitc(inducing=r.uniform(size=(300, 2)))), y, {})]
for name, gp, yy, kw in cases:
    try:
        g = G.gp_fit(gp, x, yy, **kw)
        print(name, "OK", G.gp_energy(g), flush=True)
    except L.GplError as e:
        print(name, "ERR", e.code, str(e).replace(chr(10), " "), flush=True)
"""

NAMES = ("fp64", "mixed", "laplace", "fitc")


def _run(extra):
    env = {k: v for k, v in os.environ.items() if k != "GPL_PANEL_WAIT_POLLS"}
    env.update(extra)
    r = subprocess.run([sys.executable, "-c", CODE], capture_output=True, text=True, cwd=ROOT, env=env,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-3000:]
    return dict((ln.split()[0], ln) for ln in r.stdout.splitlines() if ln.split() and ln.split()[0] in NAMES)


def test_forced_wait_timeout_is_reported_as_a_device_error():
    out = _run({"GPL_PANEL_WAIT_POLLS": "0"})
    for name in NAMES:
        assert " ERR " in out[name] and "timed out" in out[name], out


def test_default_wait_limit_fits():
    out = _run({})
    for name in NAMES:
        assert " OK " in out[name], out
