"""CPU-only checks: the C-ABI library loads and exports everything the header
declares, the host-side mirror packs parameters like the reference (and the
oracle), and the committed golden vectors are reproduced by the oracle."""

import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import gplite_amd as G
from gplite_amd import _lib, api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "gplite_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gpl_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build with python -m gplite_amd.build"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = _header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_library_has_no_cpu_fallback_without_gpu():
    """With every GPU hidden from HIP (here, and on a GPU machine alike) the
    context cannot be created and a fit raises instead of computing on the host."""
    import subprocess, sys
    code = ("import numpy as np\n"
            "import gplite_amd as G\n"
            "from gplite_amd import _lib as L\n"
            "try:\n"
            "    G.gp_fit(G.gp_init(G.cf_sexp(), G.lik_gaussian()), np.zeros((4, 1)), np.zeros(4))\n"
            "    print('COMPUTED')\n"
            "except L.GplError as e:\n"
            "    print('RAISED', e.code)\n")
    env = {**os.environ, "HIP_VISIBLE_DEVICES": "-1", "ROCR_VISIBLE_DEVICES": "-1"}
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, env=env, timeout=300)
    assert "RAISED" in r.stdout and "COMPUTED" not in r.stdout, (r.stdout, r.stderr[-2000:])


def test_grad_len_matches_layout():
    lib = _lib.lib()
    cfs = [G.cf_sexp(lscale=[0.5, 0.7, 0.9]), G.cf_const(), G.cf_nn(), G.cf_periodic(cf_base=G.cf_matern32())]
    terms, nt, xp, mask = api._program(cfs, np.zeros((5, 3)))
    assert lib.gpl_grad_len(terms, nt, _lib.LIK_GAUSSIAN) == (3 + 1) + 1 + 3 + (1 + 1 + 1) + 1
    assert lib.gpl_grad_len(terms, nt, _lib.LIK_BERNOULLI) == 11
    # column layout: sexp 3 cols, const none, nn 3 cols, periodic 3 cols
    assert [terms[k].col0 for k in range(nt)] == [0, 0, 3, 6]
    assert xp.shape == (5, 9)


CASES = [
    [G.cf_sexp(lscale=[0.7, 1.3], magn=1.2)],
    [G.cf_matern52(lscale=0.9), G.cf_const(magn=0.5)],
    [G.cf_nn(sigma0=1.0, sigma=1.5)],
    [G.cf_lin(magn=0.6), G.cf_sexp()],
    [G.cf_periodic(vars=[0], period=2.0, cf_base=G.cf_matern52(lscale=0.8))],
]


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_param_packing_matches_oracle(ci):
    cfs = CASES[ci]
    gp = G.gp_init(cfs, G.lik_gaussian(sigma=0.3))
    spec = [c.spec() for c in gp.cfs]
    p = G.get_param(gp)
    np.testing.assert_allclose(list(p.values()), O.get_param(spec, gp.lik.spec()), rtol=1e-15)
    assert len(p) == len(O.param_names(spec, gp.lik.spec()))
    assert G.lpdf_prior(gp) == pytest.approx(O.lpdf_prior(spec, gp.lik.spec()), rel=1e-13, abs=1e-15)
    q = np.random.default_rng(ci).normal(size=len(p))
    g2 = G.set_param(gp, q)
    np.testing.assert_allclose(list(G.get_param(g2).values()), q, rtol=1e-14)
    assert not g2.fitted


def test_product_and_fixed_priors():
    with pytest.warns(UserWarning):
        c = G.cf_sexp() * G.cf_matern32()
    assert isinstance(c, G.cf_prod) and len(c.cfs) == 2
    c = G.cf_sexp() * G.cf_matern32(prior_magn=G.prior_fixed())
    gp = G.gp_init(c, G.lik_gaussian())
    names = list(G.get_param(gp).keys())
    assert names == ["cf_sexp.lscale", "cf_sexp.magn", "cf_matern32.lscale", "lik_gaussian.sigma"]


def test_constructor_errors():
    with pytest.raises(ValueError):
        G.gp_init(G.cf_sexp(), G.lik_gaussian(), approx=G.approx_ep())  # R/gp.R:54-56
    with pytest.raises(ValueError):
        G.approx_ep(damping=1.0)  # R/approx.R:69-71
    with pytest.raises(ValueError):
        G.lik_bernoulli("cauchit")


def test_refit_guard_without_gpu():
    """R tests/testthat/test_pred.R:85-119: a fit invalidated by set_param raises."""
    gp = G.gp_init(G.cf_sexp(), G.lik_gaussian())
    gp.fit = object()
    gp.fitted = False
    with pytest.raises(RuntimeError):
        G.gp_energy(gp)


def test_golden_c1_reproduced_by_oracle():
    from tests.golden.make_golden import build
    ref = np.load(os.path.join(ROOT, "tests", "golden", "c1.npz"))
    new = build()
    for k in ref.files:
        np.testing.assert_allclose(new[k], ref[k], rtol=1e-12, atol=1e-14, err_msg=k)


def test_product_magnitude_warnings():
    """tests/testthat/test_warnings.R: a product of two covariance functions with
    non-fixed magnitudes warns (cf_prod and *, both orders); with one magnitude
    fixed it is silent (R/misc.R:58-81)."""
    import itertools
    import warnings as W

    free = [lambda: G.cf_const(), lambda: G.cf_lin(), lambda: G.cf_sexp(), lambda: G.cf_matern32(),
            lambda: G.cf_matern52(), lambda: G.cf_nn(), lambda: G.cf_periodic()]
    fixed = [lambda: G.cf_const(prior_magn=G.prior_fixed()), lambda: G.cf_lin(prior_magn=G.prior_fixed()),
             lambda: G.cf_sexp(prior_magn=G.prior_fixed()), lambda: G.cf_matern32(prior_magn=G.prior_fixed()),
             lambda: G.cf_matern52(prior_magn=G.prior_fixed()), lambda: G.cf_nn(prior_magn=G.prior_fixed()),
             lambda: G.cf_periodic(cf_base=G.cf_sexp(prior_magn=G.prior_fixed()))]
    for a, b in itertools.combinations(range(len(free)), 2):
        for make in (lambda p, q: G.cf_prod(p, q), lambda p, q: p * q):
            for p, q in ((free[a](), free[b]()), (free[b](), free[a]())):
                with pytest.warns(UserWarning, match="more than one non-fixed magnitude"):
                    make(p, q)
    for i in range(len(free)):
        for j in range(len(fixed)):
            for make in (lambda p, q: G.cf_prod(p, q), lambda p, q: p * q):
                with W.catch_warnings():
                    W.simplefilter("error")
                    make(free[i](), fixed[j]())
                    make(fixed[j](), free[i]())
            # products of three: adding a fixed magnitude stays silent, a free one warns
            for k in range(len(free)):
                with W.catch_warnings():
                    W.simplefilter("error")
                    tmp = free[i]() * fixed[j]()
                    G.cf_prod(tmp, fixed[k]())
                    G.cf_prod(fixed[k](), tmp)
                    tmp * fixed[j]()
                    fixed[j]() * tmp
                for make in (lambda p, q: G.cf_prod(p, q), lambda p, q: p * q):
                    with pytest.warns(UserWarning, match="more than one non-fixed magnitude"):
                        make(tmp, free[k]())
                    with pytest.warns(UserWarning, match="more than one non-fixed magnitude"):
                        make(free[k](), tmp)


def test_gp_compare():
    """R/gp_loo.R:147-175: differences against the best model (or ref), n * mean and
    n * sd / sqrt(n) of the pointwise differences; other objects are rejected."""
    r = np.random.default_rng(3)
    a = r.normal(size=20) - 1.0
    b = a + r.normal(scale=0.1, size=20) - 0.05
    la = {"loo": float(a.sum()), "loos": a}
    lb = {"loo": float(b.sum()), "loos": b}
    res = G.gp_compare(la, lb)
    best = 0 if a.sum() > b.sum() else 1
    d = [a, b][1 - best] - [a, b][best]
    assert np.allclose(res[best], [0.0, 0.0])
    assert np.isclose(res[1 - best, 0], d.sum())
    assert np.isclose(res[1 - best, 1], 20 * d.std(ddof=1) / np.sqrt(20))
    res1 = G.gp_compare(la, lb, ref=1)
    assert np.isclose(res1[0, 0], (a - b).sum())
    with pytest.raises(TypeError):
        G.gp_compare(la, [1, 2])
