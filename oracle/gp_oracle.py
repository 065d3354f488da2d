"""FP64 restatement of gplite's dense GP core (test infrastructure only).

Covariance functions are plain dicts (the reference's cf objects are R lists):

    {"kind": "const", "magn": m}
    {"kind": "lin", "magn": m, "vars": [...]|None, "normalize": bool}
    {"kind": "sexp"|"matern32"|"matern52", "lscale": float|array, "magn": m, ...}
    {"kind": "nn", "sigma0": s0, "sigma": s, "magn": m, ...}   (normalize defaults True)
    {"kind": "periodic", "period": p, "base": <stationary cf>, "vars": [...]|None}
    {"kind": "prod", "cfs": [cf, cf, ...]}

A model's covariance is a list of cfs (summed).  ``vars`` are 0-based column
indices.  ``means``/``scales`` are filled by :func:`learn_scales`.
Likelihoods: {"kind": "gaussian", "sigma"}, {"kind": "bernoulli"|"binomial",
"link": "logit"|"probit"}, {"kind": "poisson", "link": "log"}.
Priors: {"kind": "logunif"|"fixed"|"lognormal"(loc, scale)|"half_t"(df, scale)},
given per parameter name in the cf/lik dict's "priors" entry.

Parity status: every public function here is pinned by a check in
tests/test_oracle.py (closed forms, scipy LAPACK, numpy hermegauss,
brute-force loops, finite differences, fixed-point invariants).
"""

from __future__ import annotations

import copy
import math
import warnings

import numpy as np
import scipy.linalg as sla
from scipy.special import digamma, gammaln, ndtr

__all__ = [
    "JITTER", "cf_sexp_c", "cf_matern32_c", "cf_matern52_c", "cf_nn_c",
    "learn_scales", "prepare_inputmat", "eval_cf", "eval_cf_grads",
    "get_param", "set_param", "param_names", "lpdf_prior", "lpdf_prior_grad",
    "loglik", "loglik_d", "loglik_d2", "loglik_d2_exact", "loglik_d3", "response", "pseudodata_la",
    "hermite_polynomials", "gauss_hermite_points", "logsumexp_w",
    "tilted_moments", "pseudodata_ep",
    "laplace_iter", "laplace", "ep_iter", "ep", "gp_fit", "gp_energy",
    "gp_energy_grad", "gp_pred", "gp_pred_cov", "gp_draw", "gp_loo", "gaussian_logz_closed_form",
    "inv_lemma_get", "inv_lemma_solve", "fitc_fit", "fitc_ep_iter", "fitc_ep", "fitc_pred", "fitc_pred_cov",
    "fitc_train_var", "fitc_draw", "prior_draw",
]

JITTER = 1e-6  # R/gp.R:276-281 get_jitter default


# ---------------------------------------------------------------------------
# Native covariance builders: src/cf.cpp
# ---------------------------------------------------------------------------

def _inv_l2(lscale, d):
    ls = np.atleast_1d(np.asarray(lscale, dtype=np.float64))
    if ls.size == 1:
        return np.full(d, 1.0 / (ls[0] * ls[0]))
    if ls.size != d:
        raise ValueError("lscale must have length 1 or ncol(x)")
    return 1.0 / (ls * ls)


def _scaled_r2(x1, x2, lscale, chunk=1024):
    """r^2 = sum_k (x1_ik - x2_jk)^2 / l_k^2, difference-then-square as
    src/cf.cpp:37 (`sum(square(x1.row(i)-x2.row(j)))`), chunked over rows."""
    x1 = np.asarray(x1, dtype=np.float64)
    x2 = np.asarray(x2, dtype=np.float64)
    d = x1.shape[1]
    w = _inv_l2(lscale, d)
    out = np.empty((x1.shape[0], x2.shape[0]))
    for s in range(0, x1.shape[0], chunk):
        diff = x1[s:s + chunk, None, :] - x2[None, :, :]
        out[s:s + chunk] = np.einsum("ijk,ijk,k->ij", diff, diff, w)
    return out


def _diag_pairs(x1, x2):
    n = min(x1.shape[0], x2.shape[0])
    return x1[:n], x2[:n]


def cf_sexp_c(x1, x2, lscale, magn, diag_only=False):
    """magn^2 exp(-0.5 r^2 / l^2): src/cf.cpp:9-42 (formula :37-38).
    ARD extension: lscale of length ncol(x)."""
    x1 = np.atleast_2d(np.asarray(x1, float)); x2 = np.atleast_2d(np.asarray(x2, float))
    if diag_only:  # src/cf.cpp:17-25 pairs row i with row i
        a, b = _diag_pairs(x1, x2)
        r2 = np.sum((a - b) ** 2 * _inv_l2(lscale, a.shape[1]), axis=1)
        return magn * magn * np.exp(-0.5 * r2)
    r2 = _scaled_r2(x1, x2, lscale)
    return magn * magn * np.exp(-0.5 * r2)


def cf_matern32_c(x1, x2, lscale, magn, diag_only=False):
    """magn^2 (1 + sqrt3 r/l) exp(-sqrt3 r/l): src/cf.cpp:48-80 (:75-76)."""
    x1 = np.atleast_2d(np.asarray(x1, float)); x2 = np.atleast_2d(np.asarray(x2, float))
    if diag_only:
        a, b = _diag_pairs(x1, x2)
        r = np.sqrt(np.sum((a - b) ** 2 * _inv_l2(lscale, a.shape[1]), axis=1))
    else:
        r = np.sqrt(_scaled_r2(x1, x2, lscale))
    s3 = math.sqrt(3.0)
    return magn * magn * (1.0 + s3 * r) * np.exp(-s3 * r)


def cf_matern52_c(x1, x2, lscale, magn, diag_only=False):
    """magn^2 (1 + sqrt5 r/l + 5 r^2/(3 l^2)) exp(-sqrt5 r/l): src/cf.cpp:86-119."""
    x1 = np.atleast_2d(np.asarray(x1, float)); x2 = np.atleast_2d(np.asarray(x2, float))
    if diag_only:
        a, b = _diag_pairs(x1, x2)
        r = np.sqrt(np.sum((a - b) ** 2 * _inv_l2(lscale, a.shape[1]), axis=1))
    else:
        r = np.sqrt(_scaled_r2(x1, x2, lscale))
    s5 = math.sqrt(5.0)
    return magn * magn * (1.0 + s5 * r + 5.0 * r * r / 3.0) * np.exp(-s5 * r)


def cf_nn_c(x1, x2, sigma0, sigma, magn, diag_only=False):
    """Neural-network (arcsine) kernel: src/cf.cpp:125-162 (:148-158)."""
    x1 = np.atleast_2d(np.asarray(x1, float)); x2 = np.atleast_2d(np.asarray(x2, float))
    s2 = sigma * sigma
    b0 = sigma0 * sigma0 * s2
    if diag_only:
        a_, b_ = _diag_pairs(x1, x2)
        a = b0 + np.sum(a_ * b_, axis=1) * s2
        b = b0 + np.sum(a_ * a_, axis=1) * s2
        c = b0 + np.sum(b_ * b_, axis=1) * s2
    else:
        a = b0 + (x1 @ x2.T) * s2
        b = (b0 + np.sum(x1 * x1, axis=1) * s2)[:, None]
        c = (b0 + np.sum(x2 * x2, axis=1) * s2)[None, :]
    u = 2.0 * a / np.sqrt((1.0 + 2.0 * b) * (1.0 + 2.0 * c))
    return magn * magn * (2.0 / math.pi) * np.arcsin(u)


# ---------------------------------------------------------------------------
# R-side composition: R/cf.R:402-497, R/misc.R:17-31, learn_scales R/cf.R:824-855
# ---------------------------------------------------------------------------

_STATIONARY = ("sexp", "matern32", "matern52")


def learn_scales(cfs, x):
    """Column means / sd (n-1) per cf: R/cf.R:836-845; periodic skips (:852-855)."""
    x = np.atleast_2d(np.asarray(x, float))
    out = []
    for cf in cfs:
        cf = copy.deepcopy(cf)
        if cf["kind"] == "prod":
            cf["cfs"] = learn_scales(cf["cfs"], x)
        elif cf["kind"] != "periodic":
            xs = x if cf.get("vars") is None else x[:, cf["vars"]]
            cf["means"] = xs.mean(axis=0)
            cf["scales"] = xs.std(axis=0, ddof=1)
        out.append(cf)
    return out


def _normalize_default(cf):
    return cf.get("normalize", cf["kind"] == "nn")


def prepare_inputmat(cf, x):
    """Column selection then (x - means)/scales if normalize: R/misc.R:17-31."""
    x = np.atleast_2d(np.asarray(x, float))
    if cf.get("vars") is not None:
        x = x[:, cf["vars"]]
    if cf["kind"] not in ("periodic", "const") and _normalize_default(cf):
        if cf.get("means") is not None and cf.get("scales") is not None:
            x = (x - cf["means"]) / cf["scales"]
    return x


def _periodic_map(cf, x):
    w = 2.0 * math.pi / cf["period"]
    return np.hstack([np.sin(w * x), np.cos(w * x)])  # R/cf.R:483-484


def _eval_one(cf, x1, x2, diag_only=False):
    kind = cf["kind"]
    if kind == "prod":  # R/cf.R:491-497
        K = 1.0
        for c in cf["cfs"]:
            K = K * _eval_one(c, x1, x2, diag_only)
        return K
    a = prepare_inputmat(cf, x1)
    b = prepare_inputmat(cf, x2)
    if kind == "const":  # R/cf.R:413-424
        n = min(a.shape[0], b.shape[0])
        if diag_only:
            return np.full(n, cf["magn"] ** 2)
        return np.full((a.shape[0], b.shape[0]), cf["magn"] ** 2)
    if kind == "lin":  # R/cf.R:426-435
        K = cf["magn"] ** 2 * (a @ b.T)
        return np.diag(K).copy() if diag_only else K
    if kind == "sexp":
        return cf_sexp_c(a, b, cf["lscale"], cf["magn"], diag_only)
    if kind == "matern32":
        return cf_matern32_c(a, b, cf["lscale"], cf["magn"], diag_only)
    if kind == "matern52":
        return cf_matern52_c(a, b, cf["lscale"], cf["magn"], diag_only)
    if kind == "nn":
        return cf_nn_c(a, b, cf["sigma0"], cf["sigma"], cf["magn"], diag_only)
    if kind == "periodic":  # R/cf.R:478-489
        base = dict(cf["base"]); base["vars"] = None; base["normalize"] = False
        return _eval_one(base, _periodic_map(cf, a), _periodic_map(cf, b), diag_only)
    raise ValueError(kind)


def eval_cf(cfs, x1, x2, diag_only=False, cfind=None):
    """Sum over the cf list: R/cf.R:402-411."""
    idx = range(len(cfs)) if cfind is None else cfind
    K = 0.0
    for k in idx:
        K = K + _eval_one(cfs[k], x1, x2, diag_only)
    return K


# ---------------------------------------------------------------------------
# Parameters (log scale): R/cf.R:327-382, R/lik.R:145-166, R/gp.R:150-165
# ARD extension: a vector lscale contributes one log-parameter per column.
# ---------------------------------------------------------------------------

def _is_fixed(obj, name):
    pr = obj.get("priors", {}).get(name)
    return pr is not None and pr.get("kind") == "fixed"


def _own_names(cf):
    kind = cf["kind"]
    if kind in ("const", "lin"):
        return ["magn"]
    if kind in _STATIONARY:
        return ["lscale", "magn"]
    if kind == "nn":
        return ["sigma0", "sigma", "magn"]
    if kind == "periodic":
        return ["period"]
    if kind == "gaussian":
        return ["sigma"]
    if kind == "betabinom":
        return ["phi"]
    return []


def param_names(cfs, lik):
    """Names in get_param.gp order (cf list order, then lik)."""
    out = []

    def visit(cf, label):
        if cf["kind"] == "prod":
            for c in cf["cfs"]:
                visit(c, "cf_" + c["kind"])
            return
        if cf["kind"] == "periodic":
            if not _is_fixed(cf, "period"):
                out.append(f"{label}.period")
            visit_base(cf["base"], label)
            return
        visit_base(cf, label)

    def visit_base(cf, label):
        for nm in _own_names(cf):
            if _is_fixed(cf, nm):
                continue
            v = np.atleast_1d(cf[nm])
            if v.size == 1:
                out.append(f"{label}.{nm}")
            else:
                out.extend(f"{label}.{nm}[{i}]" for i in range(v.size))

    for cf in cfs:
        visit(cf, "cf_" + cf["kind"])
    for nm in _own_names(lik):
        if not _is_fixed(lik, nm):
            out.append(f"lik_{lik['kind']}.{nm}")
    return out


def _param_slots(cfs, lik):
    """List of (object, name, index-or-None) in get_param order."""
    slots = []

    def visit(cf):
        if cf["kind"] == "prod":
            for c in cf["cfs"]:
                visit(c)
            return
        if cf["kind"] == "periodic":
            if not _is_fixed(cf, "period"):
                slots.append((cf, "period", None))
            visit(cf["base"])
            return
        for nm in _own_names(cf):
            if _is_fixed(cf, nm):
                continue
            v = np.atleast_1d(cf[nm])
            if v.size == 1:
                slots.append((cf, nm, None))
            else:
                slots.extend((cf, nm, i) for i in range(v.size))

    for cf in cfs:
        visit(cf)
    for nm in _own_names(lik):
        if not _is_fixed(lik, nm):
            slots.append((lik, nm, None))
    return slots


def get_param(cfs, lik):
    """log of every free parameter: R/cf.R:334 (`param <- log(param)`)."""
    vals = []
    for obj, nm, i in _param_slots(cfs, lik):
        v = obj[nm] if i is None else np.atleast_1d(obj[nm])[i]
        vals.append(math.log(float(np.atleast_1d(v)[0] if i is None else v)))
    return np.array(vals)


def set_param(cfs, lik, p):
    """exp() back into the objects: R/cf.R:369-382, R/lik.R:160-166."""
    cfs = copy.deepcopy(cfs); lik = copy.deepcopy(lik)
    for (obj, nm, i), v in zip(_param_slots(cfs, lik), p):
        if i is None:
            obj[nm] = math.exp(v)
        else:
            arr = np.array(np.atleast_1d(obj[nm]), dtype=float)
            arr[i] = math.exp(v)
            obj[nm] = arr
    return cfs, lik


def _lpdf_one(prior, p):
    kind = (prior or {"kind": "logunif"})["kind"]
    if kind in ("logunif", "fixed"):  # R/priors.R:96-102
        return 0.0
    if kind == "lognormal":  # R/priors.R:110-112 (dnorm on the log-parameter)
        loc, sc = prior.get("loc", 0.0), prior.get("scale", 1.0)
        return -0.5 * math.log(2 * math.pi) - math.log(sc) - 0.5 * ((p - loc) / sc) ** 2
    if kind == "half_t":  # R/priors.R:104-108
        df, sc = prior.get("df", 1.0), prior.get("scale", 1.0)
        t = math.exp(p) / sc
        ldt = (math.lgamma((df + 1) / 2) - math.lgamma(df / 2) - 0.5 * math.log(df * math.pi)
               - (df + 1) / 2 * math.log1p(t * t / df))
        return math.log(2) - math.log(sc) + ldt + p
    raise ValueError(kind)


def _lpdf_grad_one(prior, p):
    kind = (prior or {"kind": "logunif"})["kind"]
    if kind in ("logunif", "fixed"):
        return 0.0
    if kind == "lognormal":
        loc, sc = prior.get("loc", 0.0), prior.get("scale", 1.0)
        return -(p - loc) / (sc * sc)
    if kind == "half_t":
        df, sc = prior.get("df", 1.0), prior.get("scale", 1.0)
        t = math.exp(p) / sc
        return 1.0 - (df + 1) * t * t / (df + t * t)
    raise ValueError(kind)


def _default_prior(obj, nm):
    if obj["kind"] == "nn" and nm in ("sigma0", "sigma"):
        return {"kind": "half_t"}  # R/cf.R:156-158 defaults
    return {"kind": "logunif"}


def lpdf_prior(cfs, lik):
    """sum of hyperprior log densities on log-params: R/gp.R:167-169, R/cf.R:503-528."""
    p = get_param(cfs, lik)
    lp = 0.0
    for (obj, nm, _i), v in zip(_param_slots(cfs, lik), p):
        lp += _lpdf_one(obj.get("priors", {}).get(nm, _default_prior(obj, nm)), v)
    return lp


def lpdf_prior_grad(cfs, lik):
    p = get_param(cfs, lik)
    return np.array([_lpdf_grad_one(obj.get("priors", {}).get(nm, _default_prior(obj, nm)), v)
                     for (obj, nm, _i), v in zip(_param_slots(cfs, lik), p)])


# ---------------------------------------------------------------------------
# dK/dlog(theta) (new work: the reference has no gradient)
# ---------------------------------------------------------------------------

def _stationary_dg(kind, rho, magn):
    """k = g(rho), rho = sum_k D_k^2/l_k^2; returns (g, dg/drho)."""
    m2 = magn * magn
    if kind == "sexp":
        g = m2 * np.exp(-0.5 * rho)
        return g, -0.5 * g
    r = np.sqrt(rho)
    if kind == "matern32":
        e = np.exp(-math.sqrt(3.0) * r)
        return m2 * (1 + math.sqrt(3.0) * r) * e, -1.5 * m2 * e
    e = np.exp(-math.sqrt(5.0) * r)
    return (m2 * (1 + math.sqrt(5.0) * r + 5.0 * rho / 3.0) * e,
            -(5.0 / 6.0) * m2 * (1 + math.sqrt(5.0) * r) * e)


def _stationary_grads(cf, a, b, dfeat_a=None, dfeat_b=None):
    """K and the list of dK/dlog(param) for a stationary kernel on features a,b.
    If dfeat_* given (periodic), also returns dK/dlog(period) first."""
    d = a.shape[1]
    w = _inv_l2(cf["lscale"], d)
    diff = a[:, None, :] - b[None, :, :]
    rho = np.einsum("ijk,ijk,k->ij", diff, diff, w)
    g, dg = _stationary_dg(cf["kind"], rho, cf["magn"])
    grads = []
    if dfeat_a is not None:
        ddiff = dfeat_a[:, None, :] - dfeat_b[None, :, :]
        grads.append(dg * 2.0 * np.einsum("ijk,ijk,k->ij", diff, ddiff, w))
    if not _is_fixed(cf, "lscale"):
        ls = np.atleast_1d(cf["lscale"])
        if ls.size == 1:
            grads.append(dg * (-2.0 * rho))
        else:
            for k in range(d):
                grads.append(dg * (-2.0 * diff[:, :, k] ** 2 * w[k]))
    if not _is_fixed(cf, "magn"):
        grads.append(2.0 * g)
    return g, grads


def _one_grads(cf, x1, x2):
    kind = cf["kind"]
    if kind == "prod":
        parts = [_one_grads(c, x1, x2) for c in cf["cfs"]]
        Ks = [p[0] for p in parts]
        K = np.prod(Ks, axis=0)
        grads = []
        for i, (_, gl) in enumerate(parts):
            others = np.prod([Ks[j] for j in range(len(Ks)) if j != i], axis=0) if len(Ks) > 1 else 1.0
            grads.extend(g * others for g in gl)
        return K, grads
    a = prepare_inputmat(cf, x1)
    b = prepare_inputmat(cf, x2)
    if kind == "const":
        K = np.full((a.shape[0], b.shape[0]), cf["magn"] ** 2)
        return K, ([] if _is_fixed(cf, "magn") else [2.0 * K])
    if kind == "lin":
        K = cf["magn"] ** 2 * (a @ b.T)
        return K, ([] if _is_fixed(cf, "magn") else [2.0 * K])
    if kind in _STATIONARY:
        return _stationary_grads(cf, a, b)
    if kind == "nn":
        s2 = cf["sigma"] ** 2
        b0 = cf["sigma0"] ** 2 * s2
        xx = a @ b.T
        aa = np.sum(a * a, 1)[:, None]
        bb = np.sum(b * b, 1)[None, :]
        A = b0 + xx * s2; Bv = b0 + aa * s2; Cv = b0 + bb * s2
        D = (1 + 2 * Bv) * (1 + 2 * Cv)
        u = 2 * A / np.sqrt(D)
        m2 = cf["magn"] ** 2
        K = m2 * (2 / math.pi) * np.arcsin(u)
        dk = m2 * (2 / math.pi) / np.sqrt(1 - u * u)
        grads = []

        def du(dA, dB, dC):
            return 2 * dA / np.sqrt(D) - u * (dB / (1 + 2 * Bv) + dC / (1 + 2 * Cv))
        if not _is_fixed(cf, "sigma0"):
            grads.append(dk * du(2 * b0, 2 * b0, 2 * b0))
        if not _is_fixed(cf, "sigma"):
            grads.append(dk * du(2 * A, 2 * Bv, 2 * Cv))
        if not _is_fixed(cf, "magn"):
            grads.append(2 * K)
        return K, grads
    if kind == "periodic":
        wp = 2 * math.pi / cf["period"]
        fa = _periodic_map(cf, a); fb = _periodic_map(cf, b)
        # d phi / d log(period) = -(w x) [cos(wx), -sin(wx)]
        da = np.hstack([-(wp * a) * np.cos(wp * a), (wp * a) * np.sin(wp * a)])
        db = np.hstack([-(wp * b) * np.cos(wp * b), (wp * b) * np.sin(wp * b)])
        base = dict(cf["base"])
        if base["kind"] not in _STATIONARY:
            raise NotImplementedError("periodic gradient needs a stationary base kernel")
        K, grads = _stationary_grads(base, fa, fb, da, db)
        if _is_fixed(cf, "period"):
            grads = grads[1:]
        return K, grads
    raise ValueError(kind)


def eval_cf_grads(cfs, x1, x2):
    """(K, [dK/dlog theta_j]) over the cf list, in get_param order."""
    K = 0.0
    grads = []
    for cf in cfs:
        Ki, gi = _one_grads(cf, x1, x2)
        K = K + Ki
        grads.extend(gi)
    return K, grads


# ---------------------------------------------------------------------------
# Likelihoods: R/lik.R
# ---------------------------------------------------------------------------

def response(lik, f):
    """R/lik.R:460-498."""
    if lik["kind"] == "gaussian":
        return f
    if lik["kind"] in ("bernoulli", "binomial", "betabinom"):
        if lik.get("link", "logit") == "probit":
            return ndtr(f)
        return 1.0 / (1.0 + np.exp(-f))
    if lik["kind"] == "poisson":
        return np.exp(f)
    raise ValueError(lik["kind"])


def _dnorm_log(x):
    return -0.5 * math.log(2 * math.pi) - 0.5 * x * x


def loglik(lik, f, y, trials=None, offset=None, total=True):
    """Observation log density: R/lik.R:297-366 (dnorm/dbinom/dpois)."""
    f = np.asarray(f, float)
    if offset is not None:
        f = f + offset
    y = np.asarray(y, float)
    kind = lik["kind"]
    if kind == "gaussian":
        s = lik["sigma"]
        ll = _dnorm_log((y - f) / s) - math.log(s)
    elif kind in ("bernoulli", "binomial"):
        t = np.ones_like(y) if kind == "bernoulli" or trials is None else np.asarray(trials, float)
        mu = response(lik, f)
        with np.errstate(divide="ignore", invalid="ignore"):
            lc = gammaln(t + 1) - gammaln(y + 1) - gammaln(t - y + 1)
            ll = lc + np.where(y > 0, y * np.log(mu), 0.0) + np.where(t - y > 0, (t - y) * np.log1p(-mu), 0.0)
    elif kind == "poisson":
        mu = response(lik, f)
        ll = np.where(y > 0, y * np.log(mu), 0.0) - mu - gammaln(y + 1)
    elif kind == "betabinom":  # R/lik.R:332-352
        t = np.asarray(trials, float)
        mu = response(lik, f)
        a = mu / lik["phi"]
        b = (1 - mu) / lik["phi"]
        t1 = gammaln(t + 1) - gammaln(y + 1) - gammaln(t - y + 1)
        t2 = gammaln(y + a) + gammaln(t - y + b) - gammaln(t + a + b)
        t3 = gammaln(a + b) - gammaln(a) - gammaln(b)
        ll = t1 + t2 + t3
    else:
        raise ValueError(kind)
    return float(np.sum(ll)) if total else ll


def loglik_d(lik, f, y, trials=None, offset=None):
    """d loglik / d f: R/lik.R:376-397 (binomial), :427-443 (poisson)."""
    f = np.asarray(f, float)
    if offset is not None:
        f = f + offset
    y = np.asarray(y, float)
    kind = lik["kind"]
    if kind in ("bernoulli", "binomial"):
        t = np.ones_like(y) if kind == "bernoulli" or trials is None else np.asarray(trials, float)
        mu = response(lik, f)
        if lik.get("link", "logit") == "probit":
            dmu = np.exp(_dnorm_log(f))
        else:
            dmu = mu * (1 - mu)
        with np.errstate(divide="ignore", invalid="ignore"):
            return dmu * (y / mu - (t - y) / (1 - mu))
    if kind == "poisson":
        mu = response(lik, f)
        return (y / mu - 1) * mu
    if kind == "betabinom":  # R/lik.R:400-425
        t = np.asarray(trials, float)
        mu = response(lik, f)
        dmu = np.exp(_dnorm_log(f)) if lik.get("link", "logit") == "probit" else mu * (1 - mu)
        phi = lik["phi"]
        a = mu / phi
        b = (1 - mu) / phi
        return (dmu / phi) * (digamma(y + a) - digamma(t - y + b)) + (dmu / phi) * (-digamma(a) + digamma(b))
    raise ValueError(kind)


def loglik_d2(lik, f, y, eps=1e-6, **kw):
    """Forward finite difference of the gradient: R/lik.R:449-454."""
    g = loglik_d(lik, f, y, **kw)
    ge = loglik_d(lik, np.asarray(f, float) + eps, y, **kw)
    return g, (ge - g) / eps


def _binom_parts(lik, f, y, trials, offset):
    f = np.asarray(f, float)
    if offset is not None:
        f = f + offset
    y = np.asarray(y, float)
    t = np.ones_like(y) if lik["kind"] == "bernoulli" or trials is None else np.asarray(trials, float)
    return f, y, t


def loglik_d2_exact(lik, f, y, trials=None, offset=None):
    """Analytic second derivative of log p(y|f) in f (not in the reference, whose
    curvature is the forward difference of R/lik.R:449-454).  logit:
    -t mu(1-mu); probit with r1 = phi/Phi, r0 = phi/(1-Phi):
    -y r1 (f + r1) - (t-y) r0 (r0 - f); Poisson (log link): -exp(f)."""
    kind = lik["kind"]
    if kind == "gaussian":
        return -np.ones(len(y)) / lik["sigma"] ** 2
    if kind in ("bernoulli", "binomial"):
        f, y, t = _binom_parts(lik, f, y, trials, offset)
        if lik.get("link", "logit") == "probit":
            phi = np.exp(_dnorm_log(f))
            r1 = phi / ndtr(f)
            r0 = phi / ndtr(-f)
            return -y * r1 * (f + r1) - (t - y) * r0 * (r0 - f)
        mu = response(lik, f)
        return -t * mu * (1 - mu)
    if kind == "poisson":
        f = np.asarray(f, float) + (0.0 if offset is None else offset)
        return -np.exp(f)
    raise NotImplementedError(f"analytic curvature for {kind}")


def loglik_d3(lik, f, y, trials=None, offset=None):
    """Third derivative of log p(y|f) in f (the implicit term of the Laplace
    gradient, Rasmussen & Williams 2006 sec. 5.5.1).  logit: -t mu(1-mu)(1-2mu);
    probit: -y r1 [1 - (f + r1)(f + 2 r1)] - (t-y) r0 [(r0 - f)(2 r0 - f) - 1];
    Poisson: -exp(f); Gaussian: 0."""
    kind = lik["kind"]
    if kind == "gaussian":
        return np.zeros(len(y))
    if kind in ("bernoulli", "binomial"):
        f, y, t = _binom_parts(lik, f, y, trials, offset)
        if lik.get("link", "logit") == "probit":
            phi = np.exp(_dnorm_log(f))
            r1 = phi / ndtr(f)
            r0 = phi / ndtr(-f)
            return -y * r1 * (1 - (f + r1) * (f + 2 * r1)) - (t - y) * r0 * ((r0 - f) * (2 * r0 - f) - 1)
        mu = response(lik, f)
        return -t * mu * (1 - mu) * (1 - 2 * mu)
    if kind == "poisson":
        f = np.asarray(f, float) + (0.0 if offset is None else offset)
        return -np.exp(f)
    raise NotImplementedError(f"third derivative for {kind}")


def pseudodata_la(lik, f, y, min_abs_curvature=1e-12, analytic_w=False, **kw):
    """Laplace pseudo-data: R/lik.R:188-201.  analytic_w replaces the reference's
    forward-difference curvature by loglik_d2_exact."""
    if lik["kind"] == "gaussian":
        n = len(y)
        return np.asarray(y, float).copy(), lik["sigma"] ** 2 * np.ones(n)
    if analytic_w:
        g, g2 = loglik_d(lik, f, y, **kw), loglik_d2_exact(lik, f, y, **kw)
    else:
        g, g2 = loglik_d2(lik, f, y, **kw)
    g2 = np.sign(g2) * np.maximum(np.abs(g2), min_abs_curvature)
    return f - g / g2, -1.0 / g2


# ---------------------------------------------------------------------------
# Quadrature: R/quadrature.R, logsumexp R/gp_loo.R:136-144
# ---------------------------------------------------------------------------

def hermite_polynomials(n):
    """Coefficients of probabilists' Hermite polynomials: R/quadrature.R:1-16."""
    c = np.zeros((n + 1, n + 1))
    c[0, 0] = 1
    if n == 0:
        return c
    c[1, 1] = 1
    for i in range(2, n + 1):
        c[i, :] = np.concatenate([[0.0], c[i - 1, :n]]) - (i - 1) * c[i - 2, :]
    return c


def gauss_hermite_points(p):
    """Nodes = roots of He_p, weights p!/(p^2 He_{p-1}(x)^2): R/quadrature.R:28-34."""
    c = hermite_polynomials(p)
    x = np.sort(np.real(np.roots(c[p, ::-1])))
    w = math.factorial(p) / (p * p * np.polyval(c[p - 1, ::-1], x) ** 2)
    return x, w


def logsumexp_w(x, w):
    m = np.max(x, axis=-1, keepdims=True)
    return (m + np.log(np.sum(w * np.exp(x - m), axis=-1, keepdims=True)))[..., 0]


def tilted_moments(lik, post_mean, post_prec, z_old, P_old, y, quad_order=11, **kw):
    """Cavity + tilted moments by GH quadrature: R/lik.R:237-270."""
    cav_prec = post_prec - P_old
    cav_mean = 1.0 / cav_prec * (post_mean * post_prec - z_old * P_old)
    gx, gw = gauss_hermite_points(quad_order)
    fgrid = cav_mean[:, None] + np.sqrt(1.0 / cav_prec)[:, None] * gx[None, :]
    n, q = fgrid.shape
    kw_rep = {k: (None if v is None else np.repeat(np.asarray(v, float), q)) for k, v in kw.items()}
    ll = loglik(lik, fgrid.ravel(), np.repeat(y, q), total=False, **kw_rep).reshape(n, q)
    lik_v = np.exp(ll)
    quad_ok = True
    if np.any(np.isnan(lik_v)):
        lik_v[np.isnan(lik_v)] = 0
        ll[np.isnan(ll)] = -np.inf
        quad_ok = False
    logC = logsumexp_w(ll, gw[None, :])
    C = np.exp(logC)
    mean_t = (1.0 / C) * ((fgrid * lik_v) @ gw)
    var_t = (1.0 / C) * (((fgrid - mean_t[:, None]) ** 2 * lik_v) @ gw)
    return mean_t, var_t, logC, cav_mean, 1.0 / cav_prec, quad_ok


def pseudodata_ep(lik, post_mean, post_prec, z_old, P_old, y, quad_order=11, damping=0.5, **kw):
    """Site update with damping: R/lik.R:207-232."""
    mt, vt, logC, cm, cv, ok = tilted_moments(lik, post_mean, post_prec, z_old, P_old, y, quad_order, **kw)
    P_new = 1 / vt - 1 / cv
    nu_new = mt / vt - cm / cv
    nu_old = z_old * P_old
    nu_new = damping * nu_new + (1 - damping) * nu_old
    P_new = damping * P_new + (1 - damping) * P_old
    return (1 / P_new) * nu_new, 1 / P_new, logC, cm, cv, ok


# ---------------------------------------------------------------------------
# Engines: R/laplace.R, R/ep.R
# ---------------------------------------------------------------------------

def _chol_B(K, V):
    """B = I + S K S, C = diag(sqrt V) chol(B)^T: R/laplace.R:27-28."""
    s = np.sqrt(1.0 / V)
    B = np.eye(len(V)) + s[:, None] * K * s[None, :]
    LB = sla.cholesky(B, lower=True)
    return np.sqrt(V)[:, None] * LB


def _sub_kw(kw, idx):
    return {k: (None if v is None else np.asarray(v, float)[idx]) for k, v in kw.items()}


def _split_sites(K, V, alpha1, C1, good, bad):
    """The bad sites' predictive N(m, P) given the good pseudo-data, P = Lk Lk^T, and
    A = I + Lk^T V_bad^-1 Lk with its factor (None if A is not positive definite):
    R/laplace.R:64-79, R/ep.R:79-92."""
    K21 = K[np.ix_(bad, good)]
    K22 = K[np.ix_(bad, bad)]
    pred_mean = K21 @ alpha1
    aux = sla.solve_triangular(C1, K21.T, lower=True)
    pred_cov = K22 - aux.T @ aux
    Lk = sla.cholesky(pred_cov, lower=True)  # R's chol() raises if this fails
    A = np.eye(len(bad)) + Lk.T @ (Lk / V[bad][:, None])
    try:
        A_chol = sla.cholesky(A, lower=True)
    except sla.LinAlgError:
        A_chol = None
    return pred_mean, pred_cov, Lk, A_chol


def _solve_via(Lk, A_chol, v):
    """backsolve(t(Lk), backsolve(t(A_chol), forwardsolve(A_chol, t(Lk) %*% v)))."""
    u = sla.solve_triangular(A_chol, Lk.T @ v, lower=True)
    u = sla.solve_triangular(A_chol, u, lower=True, trans="T")
    return sla.solve_triangular(Lk, u, lower=True, trans="T")


def laplace_iter(lik, K, y, fhat_old, pobs=None, analytic_w=False, **kw):
    """One Newton step: R/laplace.R:13-117,119 (log-concave branch, and the split
    branch for sites with pseudo-variance <= 0, with the final solve by LU as
    Matrix::lu does)."""
    if pobs is None:
        z, V = pseudodata_la(lik, fhat_old, y, analytic_w=analytic_w, **kw)
    else:
        z, V = pobs
    if not np.all(V > 0):
        good = np.flatnonzero(V > 0)
        bad = np.flatnonzero(~(V > 0))
        if len(bad) > 500:
            warnings.warn("More than 500 data points with pseudo-variance < 0 (outliers), computation getting slow. "
                          "Check that the model and hyperparameter values are sensible.")
        fit1 = laplace_iter(lik, K[np.ix_(good, good)], np.asarray(y)[good], None, pobs=(z[good], V[good]),
                            **_sub_kw(kw, good))
        pred_mean, pred_cov, Lk, A_chol = _split_sites(K, V, fit1["alpha"], fit1["C_chol"], good, bad)
        if A_chol is None:
            return dict(fmean=fhat_old, log_evidence=-np.inf)
        alpha2 = _solve_via(Lk, A_chol, z[bad] / V[bad])
        aux = sla.solve_triangular(Lk, sla.solve_triangular(Lk, pred_mean, lower=True), lower=True, trans="T")
        aux = _solve_via(Lk, A_chol, aux)
        f2hat = pred_cov @ (alpha2 + aux)
        invC2_V2_logdet = 2 * np.sum(np.log(np.diag(A_chol)))
        df2 = float(np.sum(sla.solve_triangular(Lk, f2hat - pred_mean, lower=True) ** 2))
        ll2 = loglik(lik, f2hat, np.asarray(y)[bad], **_sub_kw(kw, bad))
        le2 = ll2 - 0.5 * df2 + 0.5 * invC2_V2_logdet
        lu = sla.lu_factor(K + np.diag(V))
        alpha = sla.lu_solve(lu, z)
        fnew = K @ alpha
        return dict(fmean=fnew, z=z, V=V, C_lu=lu, alpha=alpha, log_evidence=float(fit1["log_evidence"] + le2))
    C = _chol_B(K, V)
    alpha = sla.solve_triangular(C, sla.solve_triangular(C, z, lower=True), lower=True, trans="T")
    fnew = K @ alpha
    C_logdet = 2 * np.sum(np.log(np.diag(C)))
    V_logdet = np.sum(np.log(V))
    ll = loglik(lik, fnew, y, **kw)
    logZ = ll - 0.5 * float(fnew @ alpha) - 0.5 * (C_logdet - V_logdet)
    return dict(fmean=fnew, z=z, V=V, C_chol=C, alpha=alpha, log_evidence=logZ)


def laplace(lik, K, y, maxiter=30, tol=1e-4, analytic_w=False, **kw):
    """Newton loop: R/laplace.R:274-297 (Gaussian forced to one step)."""
    n = len(y)
    fhat = np.zeros(n)
    if lik["kind"] == "gaussian":
        maxiter = 1
    it = 0
    diff = np.inf
    fit = None
    for it in range(1, maxiter + 1):
        fit = laplace_iter(lik, K, y, fhat, analytic_w=analytic_w, **kw)
        diff = float(np.max(np.abs(fhat - fit["fmean"])))
        fhat = fit["fmean"]
        if diff < tol:
            break
    if maxiter > 1 and it == maxiter:
        warnings.warn(f"Maximum number of iterations in Laplace reached, max delta f = {diff}")
    fit["iters"] = it
    return fit


def ep_iter(lik, K, y, fmean_old, fvar_old, z_old, P_old, quad_order=11, damping=0.9, pobs=None, **kw):
    """Parallel EP sweep: R/ep.R:16-146 (log-concave branch, and the split branch for
    sites with variance <= 0, with the final solves by LU as Matrix::lu does)."""
    if pobs is None:
        z, V, logZ_site, cm, cv, ok = pseudodata_ep(lik, fmean_old, 1 / fvar_old, z_old, P_old, y,
                                                    quad_order, damping, **kw)
    else:
        z, V, logZ_site, cm, cv, ok = pobs
    if np.any(np.isnan(V)) or np.any(np.isnan(z)):
        return dict(fmean=fmean_old, log_evidence=-np.inf, nan=True)
    if not np.all(V > 0):
        good = np.flatnonzero(V > 0)
        bad = np.flatnonzero(~(V > 0))
        if len(bad) > 500:
            warnings.warn("More than 500 data points with pseudo-variance < 0 (outliers), computation getting slow. "
                          "Check that the model and hyperparameter values are sensible.")
        fit1 = ep_iter(lik, K[np.ix_(good, good)], np.asarray(y)[good], None, None, None, None,
                       pobs=(z[good], V[good], logZ_site[good], cm[good], cv[good], ok), **_sub_kw(kw, good))
        pred_mean, pred_cov, Lk, A_chol = _split_sites(K, V, fit1["alpha"], fit1["C_chol"], good, bad)
        if A_chol is None:
            return dict(fmean=np.inf, log_evidence=-np.inf, no_posterior=True)
        invC2_V2_logdet = 2 * np.sum(np.log(np.diag(A_chol)))
        d = pred_mean - z[bad]
        df2 = float(d @ _solve_via(Lk, A_chol, d / V[bad]))
        le2 = (-0.5 * df2 + 0.5 * invC2_V2_logdet + np.sum(logZ_site[bad])
               + 0.5 * np.sum(np.log((cv[bad] + V[bad]) / V[bad])) + 0.5 * np.sum((cm[bad] - z[bad]) ** 2 / (cv[bad] + V[bad])))
        lu = sla.lu_factor(K + np.diag(V))
        alpha = sla.lu_solve(lu, z)
        fmean = K @ alpha
        fvar = V * np.diag(sla.lu_solve(lu, K))
        return dict(fmean=fmean, fvar=fvar, cavity_mean=cm, cavity_var=cv, z=z, P=1 / V, C_lu=lu, quad_ok=ok,
                    alpha=alpha, log_evidence=float(fit1["log_evidence"] + le2))
    C = _chol_B(K, V)
    alpha = sla.solve_triangular(C, sla.solve_triangular(C, z, lower=True), lower=True, trans="T")
    fmean = K @ alpha
    W = sla.solve_triangular(C, sla.solve_triangular(C, K, lower=True), lower=True, trans="T")
    fvar = V * np.diag(W)
    C_logdet = 2 * np.sum(np.log(np.diag(C)))
    logZ = (-0.5 * C_logdet - 0.5 * float(z @ alpha) + np.sum(logZ_site)
            + 0.5 * np.sum(np.log(cv + V)) + 0.5 * np.sum((cm - z) ** 2 / (cv + V)))
    return dict(fmean=fmean, fvar=fvar, cavity_mean=cm, cavity_var=cv, z=z, P=1 / V,
                C_chol=C, alpha=alpha, log_evidence=float(logZ), quad_ok=ok)


def ep(lik, K, y, maxiter=100, tol=1e-4, damping=0.9, damping_min=0.1, quad_order=11, **kw):
    """Damped parallel-EP loop: R/ep.R:286-330."""
    n = len(y)
    P_old = np.zeros(n); z_old = np.zeros(n)
    fmean_old = np.zeros(n); fvar_old = np.diag(K).copy()
    diff_old = np.inf
    damp = damping
    it = 0
    fit = None
    diff_new = np.inf
    for it in range(1, maxiter + 1):
        fit = ep_iter(lik, K, y, fmean_old, fvar_old, z_old, P_old, quad_order, damp, **kw)
        if fit.get("nan"):
            # R/ep.R:26-27 returns list(fmean_new = fmean_old, log_evidence = -Inf); `fit$fmean`
            # then partially matches fmean_new, so diff_new = 0 < tol and the loop ends here
            break
        if fit.get("no_posterior"):
            # R/ep.R:90-92 returns fmean = Inf: diff_new = Inf, a rejected sweep after an
            # accepted one; on the first sweep R carries Inf into the next sweep and fails
            if np.isinf(diff_old):
                raise FloatingPointError("EP: the posterior is not proper on the first sweep")
            diff_new = np.inf
            damp = damp * damping
            continue
        diff_new = float(np.max(np.abs(fmean_old - fit["fmean"])))
        if diff_new > diff_old:
            damp = damp * damping
            continue
        damp = max(damp, damping_min)
        fmean_old = fit["fmean"]; fvar_old = fit["fvar"]; z_old = fit["z"]; P_old = fit["P"]
        diff_old = diff_new
        if diff_new < tol:
            break
    if maxiter > 1 and it == maxiter:
        fit["log_evidence"] = -np.inf
        warnings.warn(f"Maximum number of iterations in EP reached, max delta f = {diff_new}")
    if fit.get("quad_ok") is False:
        fit["log_evidence"] = -np.inf
    fit["iters"] = it
    return fit


# ---------------------------------------------------------------------------
# User-level: R/gp_fit.R, R/gp.R, R/gp_pred.R
# ---------------------------------------------------------------------------

def gp_fit(cfs, lik, x, y, approx="laplace", maxiter=None, tol=1e-4, jitter=None,
           damping=0.9, quad_order=11, normalize_learn=True, analytic_w=False, ep_tol=1e-4, **kw):
    """fit_laplace/fit_ep.method_full: R/gp_fit.R:54-60,91-99,150-161.  analytic_w:
    Laplace with the analytic curvature; ep_tol: EP's stopping tolerance (the
    reference fixes 1e-4, R/ep.R:286)."""
    x = np.atleast_2d(np.asarray(x, float))
    if x.shape[0] == 1 and len(y) != 1:
        x = x.T
    if normalize_learn:
        cfs = learn_scales(cfs, x)
    jitter = JITTER if jitter is None else jitter
    n = len(y)
    K = eval_cf(cfs, x, x) + jitter * np.eye(n)
    if approx == "laplace":
        fit = laplace(lik, K, y, maxiter=30 if maxiter is None else maxiter, tol=tol, analytic_w=analytic_w, **kw)
    elif approx == "ep":
        if lik["kind"] == "gaussian":
            raise ValueError("approx_ep disabled for lik_gaussian")  # R/gp.R:54-56
        fit = ep(lik, K, y, maxiter=100 if maxiter is None else maxiter, tol=ep_tol,
                 damping=damping, damping_min=0.1, quad_order=quad_order, **kw)
    else:
        raise ValueError(approx)
    fit["cfs"] = cfs; fit["lik"] = lik; fit["x"] = x; fit["K"] = K; fit["jitter"] = jitter
    fit["approx"] = approx; fit["analytic_w"] = bool(analytic_w); fit["y"] = np.asarray(y, float)
    fit["trials"] = kw.get("trials"); fit["offset"] = kw.get("offset")
    return fit


def gp_energy(fit, include_prior=True):
    """-log evidence - log prior: R/gp.R:138-147."""
    e = -fit["log_evidence"]
    if include_prior:
        e -= lpdf_prior(fit["cfs"], fit["lik"])
    return e


def gp_energy_grad(fit, include_prior=True):
    """d energy / d log(theta) (new work, the reference has none).

    Gaussian: -1/2 tr((a a^T - Ky^-1) dKy/dtheta), Rasmussen & Williams (2006)
    eq. 5.9, with Ky = K_cf + (jitter + sigma^2) I.

    EP: the same form with the site variances in place of the noise,
    -1/2 [a^T C_j a - tr((K + V)^-1 C_j)], a = (K + V)^-1 z (R&W eq. 5.27: at
    the fixed point the site parameters do not move the evidence to first order).

    Laplace with the analytic curvature (R&W Algorithm 5.1): the explicit term
    with R = (K + W^-1)^-1, plus s2^T (I - K R) C_j grad log p, where
    s2 = d logZ / d fhat = 1/2 diag((K^-1 + W)^-1) d3 log p (the -1/2 log|B|
    term through W = -d2 log p, whose derivative in f is -d3 log p).  Only the covariance parameters enter; the
    likelihood's parameters are not differentiated for these two."""
    lik = fit["lik"]
    if lik["kind"] != "gaussian":
        return _grad_latent(fit, include_prior)
    x = fit["x"]; n = x.shape[0]
    _, dK = eval_cf_grads(fit["cfs"], x, x)
    Ky = fit["K"] + lik["sigma"] ** 2 * np.eye(n)
    L = sla.cholesky(Ky, lower=True)
    Kinv = sla.cho_solve((L, True), np.eye(n))
    a = fit["alpha"]
    g = [-0.5 * (float(a @ (D @ a)) - float(np.sum(Kinv * D))) for D in dK]
    if not _is_fixed(lik, "sigma"):
        s2 = 2 * lik["sigma"] ** 2
        g.append(-0.5 * s2 * (float(a @ a) - float(np.trace(Kinv))))
    g = np.array(g)
    if include_prior:
        g = g - lpdf_prior_grad(fit["cfs"], lik)
    return g


def _grad_latent(fit, include_prior):
    if "C_chol" not in fit:
        raise NotImplementedError("the gradient needs a fit without split sites")
    if fit["approx"] == "laplace" and not fit.get("analytic_w"):
        raise NotImplementedError("the Laplace gradient is defined for the analytic curvature")
    x = fit["x"]
    _, dK = eval_cf_grads(fit["cfs"], x, x)
    K = fit["K"]
    V = fit["V"] if "V" in fit else 1.0 / fit["P"]
    Kv = K + np.diag(V)
    Lc = sla.cholesky(Kv, lower=True)
    R = sla.cho_solve((Lc, True), np.eye(len(V)))  # (K + V)^-1 = (K + W^-1)^-1
    a = fit["alpha"]
    g = np.array([0.5 * (float(a @ (D @ a)) - float(np.sum(R * D))) for D in dK])
    if fit["approx"] == "laplace":
        # implicit term: the mode moves with theta, d fhat = (I + K W)^-1 C_j grad log p
        kw = dict(trials=fit.get("trials"), offset=fit.get("offset"))
        f = fit["fmean"]
        u = loglik_d(fit["lik"], f, fit["y"], **kw)
        d3 = loglik_d3(fit["lik"], f, fit["y"], **kw)
        dg = np.diag(K) - np.einsum("ij,ji->i", K @ R, K)  # diag((K^-1 + W)^-1) = diag(K - K R K)
        # d logZ / d fhat_i = -1/2 [(K^-1 + W)^-1]_ii dW_ii/dfhat_i, with dW/df = -d3
        s2 = 0.5 * dg * d3
        for j, D in enumerate(dK):
            b = D @ u
            g[j] += float(s2 @ (b - K @ (R @ b)))
    g = -g
    if include_prior:
        g = g - lpdf_prior_grad(fit["cfs"], fit["lik"])
    return g


def gp_pred(fit, xt, var=False, jitter=None, prior=False, cfind=None):
    """Latent predictive mean / variance: R/gp_pred.R:142-156 (prior),
    :202-232 (posterior; forms K_tt and aux^T aux as the reference does).
    cfind: 0-based indices of the covariance functions kept in K_t and K_tt
    (R/gp_pred.R:206-207); alpha and C_chol stay those of the full fit."""
    cfs = fit["cfs"]
    xt = np.atleast_2d(np.asarray(xt, float))
    if prior:
        mean = np.zeros(xt.shape[0])
        if var:
            return mean, eval_cf(cfs, xt, xt, diag_only=True, cfind=cfind)
        return mean, None
    Kt = eval_cf(cfs, xt, fit["x"], cfind=cfind)
    mean = Kt @ fit["alpha"]
    if not var:
        return mean, None
    if "C_lu" in fit:  # R/gp_pred.R:220-223: var_reduction = Kt (K + V)^-1 Kt^T
        red = np.sum(Kt.T * sla.lu_solve(fit["C_lu"], Kt.T), axis=0)
        return mean, eval_cf(cfs, xt, xt, diag_only=True, cfind=cfind) - red
    aux = sla.solve_triangular(fit["C_chol"], Kt.T, lower=True)
    v = eval_cf(cfs, xt, xt, diag_only=True, cfind=cfind) - np.sum(aux * aux, axis=0)
    return mean, v


def gp_pred_cov(fit, xt, jitter=None, cfind=None):
    """Posterior mean and full covariance + jitter I: R/gp_pred.R:202-227
    (cov = TRUE branch, K_tt - aux^T aux with aux = C_chol^-1 K_t^T)."""
    cfs = fit["cfs"]
    xt = np.atleast_2d(np.asarray(xt, float))
    jitter = JITTER if jitter is None else jitter
    Kt = eval_cf(cfs, xt, fit["x"], cfind=cfind)
    mean = Kt @ fit["alpha"]
    aux = sla.solve_triangular(fit["C_chol"], Kt.T, lower=True)
    cov = eval_cf(cfs, xt, xt, cfind=cfind) - aux.T @ aux + jitter * np.eye(xt.shape[0])
    return mean, cov


def gp_draw(fit, xt, z, jitter=None, cfind=None):
    """Joint posterior draws of the latent function: R/gp_draw.R:101-118 with
    mvnrnd (R/gp_draw.R:182-200): mean + t(chol(cov)) z for the given standard
    normal matrix z of shape (len(xt), draws)."""
    mean, cov = gp_pred_cov(fit, xt, jitter, cfind)
    Lc = sla.cholesky(cov, lower=True)
    return mean[:, None] + Lc @ np.asarray(z, float).reshape(len(mean), -1)


def _logsumexp_rows(a, w=None):
    """log(sum_j w_j exp(a_ij)) per row, stabilised by the row max (R/gp_loo.R:136-144)."""
    c = np.max(a, axis=1, keepdims=True)
    e = np.exp(a - c)
    if w is not None:
        e = e * w[None, :]
    return c[:, 0] + np.log(np.sum(e, axis=1))


def gp_loo(fit, lik, x, y, quadrature=True, quad_order=11, draws=4000, seed=0, trials=None, offset=None):
    """LOO predictive log densities for method_full: R/gp_loo.R:73-110.
    Laplace LOO posteriors from the posterior at the training inputs
    (R/gp_loo.R:120-130); EP uses the cavity moments (:132-134)."""
    y = np.asarray(y, float)
    n = len(y)
    if "cavity_mean" in fit and fit.get("approx") == "ep":
        m, v = fit["cavity_mean"], fit["cavity_var"]
    else:
        fhat, z, V = fit["fmean"], fit["z"], fit["V"]
        grad = (z - fhat) / V
        if "inducing" in fit:  # method_fitc: train = TRUE
            post_var = fitc_train_var(fit, x)
        else:
            _, post_var = gp_pred(fit, x, var=True)
        v = 1.0 / (1.0 / post_var - 1.0 / V)
        m = fhat - v * grad
    sd = np.sqrt(v)
    if quadrature:
        xg, wg = gauss_hermite_points(quad_order)
        fgrid = m[:, None] + sd[:, None] * xg[None, :]
        ll = loglik(lik, fgrid, y[:, None], None if trials is None else np.asarray(trials, float)[:, None],
                    None if offset is None else np.asarray(offset, float)[:, None], total=False)
        loos = _logsumexp_rows(ll, wg / np.sum(wg))
    else:
        rng = np.random.default_rng(seed)
        fs = m[:, None] + sd[:, None] * rng.standard_normal((n, draws))
        ll = loglik(lik, fs, y[:, None], None if trials is None else np.asarray(trials, float)[:, None],
                    None if offset is None else np.asarray(offset, float)[:, None], total=False)
        loos = _logsumexp_rows(ll) - math.log(draws)
    return {"loo": float(np.sum(loos)), "sd": float(n * np.std(loos, ddof=1) / math.sqrt(n)), "loos": loos,
            "loo_mean": m, "loo_var": v}


# ---------------------------------------------------------------------------
# FITC (method_fitc): R/gp_fit.R:101-123, R/laplace.R:122-152,300-322,
# R/linear_algebra.R:3-172, R/gp_pred.R:234-290
# ---------------------------------------------------------------------------

def inv_lemma_get(K, U, D, logdet=True):
    """inv(U K^-1 U^T + diag(D)) by the inversion lemma: R/linear_algebra.R:3-47
    (V = t(U), symmetrised V D^-1 U)."""
    D = np.asarray(D, float)
    inv_D_U = U / D[:, None]
    VDU = U.T @ inv_D_U
    VDU = 0.5 * (VDU + VDU.T)
    L = sla.cholesky(K + VDU, lower=True)
    ld = None
    if logdet:
        ld = float(np.sum(np.log(D)) - 2 * np.sum(np.log(np.diag(sla.cholesky(K, lower=True))))
                   + 2 * np.sum(np.log(np.diag(L))))
    return {"chol": L, "D": D, "inv_D_U": inv_D_U, "logdet": ld}


def inv_lemma_solve(obj, rhs, lhs=None, rhs_diag=False, lhs_diag=False, diag_only=False):
    """R/linear_algebra.R:49-172: lhs C^-1 rhs with C^-1 from inv_lemma_get; a
    diagonal rhs or lhs is given as its vector, diag_only (both diagonal) returns
    the diagonal of the product."""
    D, IDU, L = obj["D"], obj["inv_D_U"], obj["chol"]
    rhs = np.asarray(rhs, float)
    if rhs_diag:
        core = sla.cho_solve((L, True), IDU.T * rhs[None, :])  # diag_times_dense(t(inv_D_U), rhs, diag_right)
        if lhs is None:
            return np.diag(rhs / D) - IDU @ core
        lhs = np.asarray(lhs, float)
        if lhs_diag:
            if diag_only:
                return lhs / D * rhs - np.sum((lhs[:, None] * IDU).T * core, axis=0)
            return np.diag(lhs / D * rhs) - (lhs[:, None] * IDU) @ core
        return lhs * (rhs / D)[None, :] - (lhs @ IDU) @ core
    core = sla.cho_solve((L, True), IDU.T @ rhs)  # backsolve(t(L), forwardsolve(L, aux_right))
    if lhs is None:
        scaled = rhs / D[:, None] if rhs.ndim == 2 else rhs / D
        return scaled - IDU @ core
    lhs = np.asarray(lhs, float)
    if lhs_diag:
        return (lhs / D)[:, None] * rhs - (lhs[:, None] * IDU) @ core
    return lhs @ (rhs / D[:, None]) - (lhs @ IDU) @ core


def fitc_ep_iter(lik, Kz, Kz_chol, Kxz, D, y, fmean_old, fvar_old, z_old, P_old, quad_order=11, damping=0.9,
                 **kw):
    """ep_iter.method_fitc, log-concave branch: R/ep.R:148-185,271-278."""
    z, V, logZ_site, cm, cv, ok = pseudodata_ep(lik, fmean_old, 1 / fvar_old, z_old, P_old, y,
                                                quad_order, damping, **kw)
    if np.any(np.isnan(V)) or np.any(np.isnan(z)):
        return dict(fmean=fmean_old, log_evidence=-np.inf, nan=True)
    if not np.all(V > 0):
        raise NotImplementedError("non-log-concave EP branch (R/ep.R:186-270) is a host fallback")
    S = D + V
    C_inv = inv_lemma_get(Kz, Kxz, S)
    alpha = inv_lemma_solve(C_inv, z)
    inv_Kz_Kzx = sla.cho_solve((Kz_chol, True), Kxz.T)
    aux = inv_lemma_solve(C_inv, V, inv_Kz_Kzx, rhs_diag=True)
    diag1 = np.sum(Kxz.T * aux, axis=0)
    diag2 = inv_lemma_solve(C_inv, V, D, rhs_diag=True, lhs_diag=True, diag_only=True)
    fmean = Kxz @ sla.cho_solve((Kz_chol, True), Kxz.T @ alpha) + D * alpha
    logZ = (-0.5 * C_inv["logdet"] - 0.5 * float(z @ alpha) + np.sum(logZ_site)
            + 0.5 * np.sum(np.log(cv + V)) + 0.5 * np.sum((cm - z) ** 2 / (cv + V)))
    return dict(fmean=fmean, fvar=diag1 + diag2, cavity_mean=cm, cavity_var=cv, z=z, V=V, P=1 / V, Kz=Kz,
                Kxz=Kxz, Kz_chol=Kz_chol, C_inv=C_inv, D=D, alpha=alpha, log_evidence=float(logZ), quad_ok=ok)


def fitc_ep(lik, Kz, Kz_chol, Kxz, K_diag, D, y, damping=0.9, damping_min=0.1, maxiter=100, tol=1e-3,
            quad_order=11, **kw):
    """ep.method_fitc: R/ep.R:332-370 (its tolerance defaults to 1e-3 and
    fit_ep.method_fitc, R/gp_fit.R:163-187, passes none)."""
    n = len(y)
    P_old = np.zeros(n); z_old = np.zeros(n)
    fmean_old = np.zeros(n); fvar_old = np.asarray(K_diag, float).copy()
    diff_old = np.inf
    damp = damping
    it = 0
    fit = None
    diff_new = np.inf
    for it in range(1, maxiter + 1):
        fit = fitc_ep_iter(lik, Kz, Kz_chol, Kxz, D, y, fmean_old, fvar_old, z_old, P_old, quad_order, damp, **kw)
        diff_new = float(np.max(np.abs(fmean_old - fit["fmean"])))
        if diff_new > diff_old:
            damp = damp * damping
            continue
        damp = max(damp, damping_min)
        fmean_old = fit["fmean"]; fvar_old = fit["fvar"]; z_old = fit["z"]; P_old = fit["P"]
        diff_old = diff_new
        if diff_new < tol:
            break
    if maxiter > 1 and it == maxiter:
        fit["log_evidence"] = -np.inf
        warnings.warn(f"Maximum number of iterations in EP reached, max delta f = {diff_new}")
    if fit.get("quad_ok") is False:
        fit["log_evidence"] = -np.inf
    fit["iters"] = it
    return fit


def fitc_fit(cfs, lik, x, y, z, maxiter=30, tol=1e-4, jitter=None, normalize_learn=True, approx="laplace",
             damping=0.9, quad_order=11, **kw):
    """fit_laplace.method_fitc (R/gp_fit.R:101-123) for given inducing inputs z,
    Newton loop of laplace.method_fitc (R/laplace.R:300-322, stopping on the
    change of the unnormalised log posterior) over the log-concave branch of
    laplace_iter.method_fitc (R/laplace.R:122-152)."""
    x = np.atleast_2d(np.asarray(x, float))
    if x.shape[0] == 1 and len(y) != 1:
        x = x.T
    z = np.atleast_2d(np.asarray(z, float))
    if normalize_learn:
        cfs = learn_scales(cfs, x)
    jitter = JITTER if jitter is None else jitter
    y = np.asarray(y, float)
    n, m = len(y), z.shape[0]
    Kz = eval_cf(cfs, z, z) + jitter * np.eye(m)
    Kxz = eval_cf(cfs, x, z)
    Kz_chol = sla.cholesky(Kz, lower=True)
    D = eval_cf(cfs, x, x, diag_only=True) + jitter
    D = D - np.sum(sla.solve_triangular(Kz_chol, Kxz.T, lower=True) ** 2, axis=0)
    if approx == "ep":  # fit_ep.method_fitc, R/gp_fit.R:163-187
        if lik["kind"] == "gaussian":
            raise ValueError("approx_ep disabled for lik_gaussian")
        K_diag = eval_cf(cfs, x, x, diag_only=True) + jitter
        fit = fitc_ep(lik, Kz, Kz_chol, Kxz, K_diag, D, y, damping=damping, damping_min=0.1,
                      maxiter=100 if maxiter is None else maxiter, quad_order=quad_order, **kw)
        fit.update(cfs=cfs, lik=lik, x=x, inducing=z, jitter=jitter, approx="ep")
        return fit
    fhat = np.zeros(n)
    if lik["kind"] == "gaussian":
        maxiter = 1
    log_post_old = -np.inf
    fit = None
    it = 0
    for it in range(1, maxiter + 1):
        pz, V = pseudodata_la(lik, fhat, y, **kw)
        S = D + V
        C_inv = inv_lemma_get(Kz, Kxz, S)
        alpha = inv_lemma_solve(C_inv, pz)
        fnew = Kxz @ sla.cho_solve((Kz_chol, True), Kxz.T @ alpha) + D * alpha
        f_invK_f = float(fnew @ alpha)
        ll = loglik(lik, fnew, y, **kw)
        logZ = ll - 0.5 * f_invK_f + 0.5 * (float(np.sum(np.log(V))) - C_inv["logdet"])
        log_post = ll - 0.5 * f_invK_f
        fit = dict(fmean=fnew, z=pz, V=V, Kz=Kz, Kxz=Kxz, Kz_chol=Kz_chol, C_inv=C_inv, D=D, alpha=alpha,
                   log_evidence=float(logZ), log_post_unnorm=float(log_post))
        fhat = fnew
        diff = log_post - log_post_old
        log_post_old = log_post
        if abs(diff) < tol:
            break
    if maxiter > 1 and it == maxiter:
        warnings.warn("Maximum number of iterations in Laplace reached")
    fit.update(iters=it, cfs=cfs, lik=lik, x=x, inducing=z, jitter=jitter, approx="laplace")
    return fit


def fitc_pred(fit, xt, var=False):
    """gp_pred_post.method_fitc with train = FALSE: R/gp_pred.R:234-290."""
    cfs, z = fit["cfs"], fit["inducing"]
    xt = np.atleast_2d(np.asarray(xt, float))
    Ktz = eval_cf(cfs, xt, z)
    Kz_chol, Kxz = fit["Kz_chol"], fit["Kxz"]
    aux1 = sla.solve_triangular(Kz_chol, Kxz.T, lower=True)
    aux2 = sla.solve_triangular(Kz_chol, Ktz.T, lower=True)
    mean = aux2.T @ (aux1 @ fit["alpha"])
    if not var:
        return mean, None
    Dt = eval_cf(cfs, xt, xt, diag_only=True) - np.sum(aux2 ** 2, axis=0)
    Kz_inv_Ktz = sla.cho_solve((Kz_chol, True), Ktz.T)
    prior_var = np.sum(Ktz * Kz_inv_Ktz.T, axis=1) + Dt
    aux = inv_lemma_solve(fit["C_inv"], aux1.T, aux1) @ aux2
    return mean, prior_var - np.sum(aux2 * aux, axis=0)


def fitc_train_var(fit, x):
    """Predictive variance at the training inputs with the FITC diagonal
    (gp_pred_post.method_fitc with train = TRUE, R/gp_pred.R:270-285), as
    loo_posteriors.approx_laplace uses it (R/gp_loo.R:120-130)."""
    cfs, z = fit["cfs"], fit["inducing"]
    x = np.atleast_2d(np.asarray(x, float))
    Ktz = eval_cf(cfs, x, z)
    Kz_chol, Kxz, C_inv = fit["Kz_chol"], fit["Kxz"], fit["C_inv"]
    aux1 = sla.solve_triangular(Kz_chol, Kxz.T, lower=True)
    aux2 = sla.solve_triangular(Kz_chol, Ktz.T, lower=True)
    Dt = eval_cf(cfs, x, x, diag_only=True) - np.sum(aux2 ** 2, axis=0)
    Kz_inv_Ktz = sla.cho_solve((Kz_chol, True), Ktz.T)
    prior_var = np.sum(Ktz * Kz_inv_Ktz.T, axis=1) + Dt
    aux = inv_lemma_solve(C_inv, aux1.T, aux1) @ aux2 + inv_lemma_solve(C_inv, Dt, aux1, rhs_diag=True)
    diag1 = np.sum(aux2 * aux, axis=0)
    diag2 = np.sum(inv_lemma_solve(C_inv, aux1.T, Dt, lhs_diag=True) * aux2.T, axis=1)
    diag3 = inv_lemma_solve(C_inv, Dt, Dt, rhs_diag=True, lhs_diag=True, diag_only=True)
    return prior_var - (diag1 + diag2 + diag3)


def fitc_pred_cov(fit, xt, jitter=None):
    """Mean and full predictive covariance + jitter I of a FITC fit, train = FALSE:
    R/gp_pred.R:258-268 (prior_cov - A_t^T A C^-1 A^T A_t)."""
    cfs, z = fit["cfs"], fit["inducing"]
    xt = np.atleast_2d(np.asarray(xt, float))
    jitter = JITTER if jitter is None else jitter
    Ktz = eval_cf(cfs, xt, z)
    Kz_chol, Kxz = fit["Kz_chol"], fit["Kxz"]
    aux1 = sla.solve_triangular(Kz_chol, Kxz.T, lower=True)
    aux2 = sla.solve_triangular(Kz_chol, Ktz.T, lower=True)
    mean = aux2.T @ (aux1 @ fit["alpha"])
    Dt = eval_cf(cfs, xt, xt, diag_only=True) - np.sum(aux2 ** 2, axis=0)
    prior_cov = aux2.T @ aux2 + np.diag(Dt)
    cov = prior_cov - aux2.T @ (aux1 @ inv_lemma_solve(fit["C_inv"], aux1.T)) @ aux2
    return mean, cov + jitter * np.eye(xt.shape[0])


def fitc_draw(fit, xt, z, jitter=None):
    """FITC joint draws, R/gp_draw.R:126-140 with mvnrnd (:182-200)."""
    mean, cov = fitc_pred_cov(fit, xt, jitter)
    Lc = sla.cholesky(cov, lower=True)
    return mean[:, None] + Lc @ np.asarray(z, float).reshape(len(mean), -1)


def prior_draw(cfs, xt, z, jitter=None):
    """Prior draws of method_full: R/gp_draw.R:56-67 with gp_pred_prior(cov = TRUE),
    R/gp_pred.R:142-151: t(chol(K_tt + jitter I)) z (mean 0)."""
    xt = np.atleast_2d(np.asarray(xt, float))
    jitter = JITTER if jitter is None else jitter
    Lc = sla.cholesky(eval_cf(cfs, xt, xt) + jitter * np.eye(xt.shape[0]), lower=True)
    return Lc @ np.asarray(z, float).reshape(xt.shape[0], -1)


def gaussian_logz_closed_form(K_cf, y, sigma, jitter=JITTER):
    """log N(y | 0, K + (sigma^2 + jitter) I) by a direct Cholesky of K_y
    (the closed form tests/testthat/test_basic.R:85-89 checks against)."""
    n = len(y)
    L = sla.cholesky(K_cf + (sigma ** 2 + jitter) * np.eye(n), lower=True)
    a = sla.solve_triangular(L, y, lower=True)
    return -0.5 * float(a @ a) - float(np.sum(np.log(np.diag(L)))) - 0.5 * n * math.log(2 * math.pi)
