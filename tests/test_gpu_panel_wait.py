"""The bounded waits between the workgroups of the single-launch Cholesky panel.

With the poll limit forced to 0 (GPL_PANEL_WAIT_POLLS, read when a context is
created), the workgroups below a diagonal block give up their first wait. Every
workgroup then leaves the kernel, and the fit reports a device error instead of a
result, in FP64 and in the mixed form. The same fits in a fresh process with the
default limit succeed. Each case runs in a child process of its own, so the
override never reaches the other tests' contexts.
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
for prec in ("fp64", "mixed"):
    gp = G.gp_init(G.cf_sexp(lscale=[0.3, 0.3]), G.lik_gaussian(sigma=0.1))
    try:
        g = G.gp_fit(gp, x, y, precision=prec)
        print(prec, "OK", G.gp_energy(g), flush=True)
    except L.GplError as e:
        print(prec, "ERR", e.code, str(e).replace(chr(10), " "), flush=True)
"""


def _run(extra):
    env = {k: v for k, v in os.environ.items() if k != "GPL_PANEL_WAIT_POLLS"}
    env.update(extra)
    r = subprocess.run([sys.executable, "-c", CODE], capture_output=True, text=True, cwd=ROOT, env=env,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-3000:]
    return dict((ln.split()[0], ln) for ln in r.stdout.splitlines() if ln.split() and ln.split()[0] in
                ("fp64", "mixed"))


def test_forced_wait_timeout_is_reported_as_a_device_error():
    out = _run({"GPL_PANEL_WAIT_POLLS": "0"})
    for prec in ("fp64", "mixed"):
        assert " ERR " in out[prec] and "timed out" in out[prec], out


def test_default_wait_limit_fits():
    out = _run({})
    for prec in ("fp64", "mixed"):
        assert " OK " in out[prec], out
