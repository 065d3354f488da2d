"""Host-side mirror of gplite's R API (gp_init / gp_fit / gp_energy / gp_pred /
gp_optim with cf_* / lik_* / prior_* / approx_* objects) over the C ABI.

The R package keeps the model as a list and dispatches on S3 classes; here the
objects are small Python classes with the same constructor arguments and
defaults.  Everything O(n^2) or worse runs in libgplite_amd.so on the GPU;
this module only does what the reference does in R at O(n*d): input column
selection and normalisation (R/misc.R:17-31), parameter packing on the log
scale (R/cf.R:327-382), hyperprior densities (R/priors.R:96-112), quantiles and
response transforms of the predictive distribution (R/gp_pred.R:102-120).
"""

from __future__ import annotations

import copy
import math
import warnings
import weakref
from ctypes import byref as C_byref, c_double as C_double, c_int32 as C_int32

import numpy as np

from . import _lib as L

# ---------------------------------------------------------------------------
# priors (R/priors.R:56-112)
# ---------------------------------------------------------------------------


class Prior:
    kind = "prior"

    def lpdf(self, p):
        return 0.0

    def dlpdf(self, p):
        return 0.0

    def spec(self):
        return {"kind": self.kind, **{k: v for k, v in self.__dict__.items()}}


class prior_fixed(Prior):
    kind = "fixed"


class prior_logunif(Prior):
    kind = "logunif"


class prior_lognormal(Prior):
    kind = "lognormal"

    def __init__(self, loc=0.0, scale=1.0):
        self.loc, self.scale = loc, scale

    def lpdf(self, p):  # dnorm on the log-parameter (R/priors.R:110-112)
        return -0.5 * math.log(2 * math.pi) - math.log(self.scale) - 0.5 * ((p - self.loc) / self.scale) ** 2

    def dlpdf(self, p):
        return -(p - self.loc) / self.scale ** 2


class prior_half_t(Prior):
    kind = "half_t"

    def __init__(self, df=1.0, scale=1.0):
        self.df, self.scale = df, scale

    def lpdf(self, p):  # R/priors.R:104-108
        t = math.exp(p) / self.scale
        df = self.df
        ldt = (math.lgamma((df + 1) / 2) - math.lgamma(df / 2) - 0.5 * math.log(df * math.pi)
               - (df + 1) / 2 * math.log1p(t * t / df))
        return math.log(2) - math.log(self.scale) + ldt + p

    def dlpdf(self, p):
        t = math.exp(p) / self.scale
        return 1.0 - (self.df + 1) * t * t / (self.df + t * t)


def _fixed(pr):
    return isinstance(pr, prior_fixed)


# ---------------------------------------------------------------------------
# covariance functions (R/cf.R:92-213)
# ---------------------------------------------------------------------------

class Cf:
    kind = "cf"
    param_names = ()

    def __init__(self):
        self.vars = None
        self.normalize = False
        self.means = None
        self.scales = None
        self.priors = {}

    # R/misc.R:17-31
    def prepare(self, x):
        x = _as_matrix(x)
        if self.vars is not None:
            x = x[:, _col_index(x, self.vars)]
        if self.normalize and self.means is not None and self.scales is not None:
            x = (x - self.means) / self.scales
        return x

    def learn_scales(self, x):  # R/cf.R:836-845
        x = _as_matrix(x)
        if self.vars is not None:
            x = x[:, _col_index(x, self.vars)]
        self.means = x.mean(axis=0)
        self.scales = x.std(axis=0, ddof=1) if x.shape[0] > 1 else np.ones(x.shape[1])

    def __mul__(self, other):  # "*.cf" (R/cf.R:193-209)
        a = self.cfs if isinstance(self, cf_prod) else [self]
        b = other.cfs if isinstance(other, cf_prod) else [other]
        return cf_prod(*a, *b)

    # parameter slots: (object, name, index or None) in get_param order
    def slots(self):
        out = []
        for nm in self.param_names:
            if _fixed(self.priors.get(nm)):
                continue
            v = np.atleast_1d(getattr(self, nm))
            if v.size == 1:
                out.append((self, nm, None))
            else:
                out.extend((self, nm, i) for i in range(v.size))
        return out

    def all_slots(self):
        out = []
        for nm in self.param_names:
            v = np.atleast_1d(getattr(self, nm))
            if v.size == 1:
                out.append((self, nm, None))
            else:
                out.extend((self, nm, i) for i in range(v.size))
        return out

    def label(self):
        return "cf_" + self.kind

    def spec(self):
        d = {"kind": self.kind, "vars": None if self.vars is None else list(self.vars),
             "normalize": self.normalize, "means": self.means, "scales": self.scales,
             "priors": {k: v.spec() for k, v in self.priors.items()}}
        for nm in self.param_names:
            d[nm] = copy.deepcopy(getattr(self, nm))
        return d


class cf_const(Cf):
    kind = "const"
    param_names = ("magn",)

    def __init__(self, magn=1.0, prior_magn=None):
        super().__init__()
        self.magn = magn
        self.priors = {"magn": prior_magn or prior_logunif()}

    def prepare(self, x):
        return _as_matrix(x)[:, :0]

    def learn_scales(self, x):
        pass


class cf_lin(Cf):
    kind = "lin"
    param_names = ("magn",)

    def __init__(self, vars=None, magn=1.0, prior_magn=None, normalize=False):
        super().__init__()
        self.vars, self.magn, self.normalize = vars, magn, normalize
        self.priors = {"magn": prior_magn or prior_logunif()}


class _Stationary(Cf):
    param_names = ("lscale", "magn")

    def __init__(self, vars=None, lscale=0.3, magn=1.0, prior_lscale=None, prior_magn=None, normalize=False):
        super().__init__()
        self.vars, self.magn, self.normalize = vars, magn, normalize
        ls = np.atleast_1d(np.asarray(lscale, dtype=float))
        self.lscale = float(ls[0]) if ls.size == 1 else ls
        self.priors = {"lscale": prior_lscale or prior_logunif(), "magn": prior_magn or prior_logunif()}


class cf_sexp(_Stationary):
    kind = "sexp"


class cf_matern32(_Stationary):
    kind = "matern32"


class cf_matern52(_Stationary):
    kind = "matern52"


class cf_nn(Cf):
    kind = "nn"
    param_names = ("sigma0", "sigma", "magn")

    def __init__(self, vars=None, sigma0=1.0, sigma=3.0, magn=1.0, prior_sigma0=None, prior_sigma=None,
                 prior_magn=None, normalize=True):
        super().__init__()
        self.vars, self.sigma0, self.sigma, self.magn, self.normalize = vars, sigma0, sigma, magn, normalize
        self.priors = {"sigma0": prior_sigma0 or prior_half_t(), "sigma": prior_sigma or prior_half_t(),
                       "magn": prior_magn or prior_logunif()}


class cf_periodic(Cf):
    kind = "periodic"
    param_names = ("period",)

    def __init__(self, vars=None, period=1.0, cf_base=None, prior_period=None):
        super().__init__()
        self.vars, self.period = vars, period
        base = copy.deepcopy(cf_base) if cf_base is not None else cf_sexp()
        if not isinstance(base, (cf_sexp, cf_matern32, cf_matern52)):
            raise ValueError("cf_periodic: base kernel must be cf_sexp, cf_matern32 or cf_matern52")
        base.normalize = False  # R/cf.R:172-173
        base.vars = None
        self.base = base
        self.priors = {"period": prior_period or prior_logunif()}

    def learn_scales(self, x):  # periodic does not normalise (R/cf.R:852-855)
        pass

    def prepare(self, x):
        x = _as_matrix(x)
        if self.vars is not None:
            x = x[:, _col_index(x, self.vars)]
        return x

    def slots(self):
        out = [] if _fixed(self.priors.get("period")) else [(self, "period", None)]
        return out + self.base.slots()

    def all_slots(self):
        return [(self, "period", None)] + self.base.all_slots()

    def spec(self):
        d = super().spec()
        d["base"] = self.base.spec()
        return d


class cf_prod(Cf):
    kind = "prod"

    def __init__(self, *cfs):
        super().__init__()
        self.cfs = list(cfs)
        # free parameters named magn, the base kernel of cf_periodic included (R/misc.R:58-81)
        nfree = sum(1 for c in self.cfs for (_o, nm, _i) in c.slots() if nm == "magn")
        if nfree > 1:
            warnings.warn("Detected a product of covariance functions with more than one non-fixed magnitude "
                          "parameters. Consider setting prior_magn = prior_fixed() for all except one of the "
                          "kernels in the product.")

    def learn_scales(self, x):
        for c in self.cfs:
            c.learn_scales(x)

    def slots(self):
        return [s for c in self.cfs for s in c.slots()]

    def all_slots(self):
        return [s for c in self.cfs for s in c.all_slots()]

    def spec(self):
        return {"kind": "prod", "cfs": [c.spec() for c in self.cfs]}


def _as_matrix(x):
    if hasattr(x, "to_numpy"):
        x = x.to_numpy()
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    return x


def _col_index(x, vars):
    return [int(v) for v in vars]


# ---------------------------------------------------------------------------
# likelihoods (R/lik.R:46-90), approximations (R/approx.R:60-75), method
# ---------------------------------------------------------------------------

class Lik:
    kind = "lik"
    param_names = ()

    def __init__(self):
        self.priors = {}
        self.link = "identity"

    def slots(self):
        return [(self, nm, None) for nm in self.param_names if not _fixed(self.priors.get(nm))]

    def all_slots(self):
        return [(self, nm, None) for nm in self.param_names]

    def label(self):
        return "lik_" + self.kind

    def spec(self):
        d = {"kind": self.kind, "link": self.link, "priors": {k: v.spec() for k, v in self.priors.items()}}
        for nm in self.param_names:
            d[nm] = getattr(self, nm)
        return d

    def response(self, f):
        f = np.asarray(f, dtype=float)
        if self.link == "logit":
            return 1.0 / (1.0 + np.exp(-f))
        if self.link == "probit":
            from scipy.special import ndtr
            return ndtr(f)
        if self.link == "log":
            return np.exp(f)
        return f


class lik_gaussian(Lik):
    kind = "gaussian"
    param_names = ("sigma",)

    def __init__(self, sigma=0.5, prior_sigma=None):
        super().__init__()
        self.sigma = sigma
        self.priors = {"sigma": prior_sigma or prior_logunif()}


class lik_bernoulli(Lik):
    kind = "bernoulli"

    def __init__(self, link="logit"):
        super().__init__()
        if link not in ("logit", "probit"):
            raise ValueError("Unknown link: " + str(link))
        self.link = link


class lik_binomial(lik_bernoulli):
    kind = "binomial"


class lik_betabinom(Lik):
    """lik_betabinom(link, phi, prior_phi) (R/lik.R:74-80): counts y of `trials` with a
    beta-distributed success probability of mean mu = link^-1(f) and overdispersion phi."""
    kind = "betabinom"
    param_names = ("phi",)

    def __init__(self, link="logit", phi=1.0, prior_phi=None):
        super().__init__()
        if link not in ("logit", "probit"):
            raise ValueError("Unknown link: " + str(link))
        self.link = link
        self.phi = phi
        self.priors = {"phi": prior_phi or prior_logunif()}


class lik_poisson(Lik):
    kind = "poisson"

    def __init__(self, link="log"):
        super().__init__()
        if link != "log":
            raise ValueError("Unknown link: " + str(link))
        self.link = link


class approx_laplace:
    """approx_laplace(maxiter, tol) of R/approx.R.  analytic_w (new, default off):
    the Newton curvature in closed form instead of the reference's forward
    difference (R/lik.R:449-454); gp_energy_grad of a Laplace fit needs it."""
    name = "laplace"

    def __init__(self, maxiter=30, tol=1e-4, analytic_w=False):
        self.maxiter, self.tol, self.analytic_w = maxiter, tol, bool(analytic_w)


class approx_ep:
    name = "ep"

    def __init__(self, damping=0.9, quad_order=11, maxiter=100):
        if damping <= 0 or damping >= 1:
            raise ValueError("Damping should be between 0 and 1 (exclusive)")
        self.damping, self.quad_order, self.maxiter = damping, quad_order, maxiter


class method_full:
    name = "full"


class method_fitc:
    """FITC with inducing inputs (R/method.R:113-124).  Without `inducing`, gp_fit
    places num_inducing points by k-means on the scaled inputs, optionally per bin
    of one input (R/gp_fit.R:195-240)."""
    name = "fitc"

    def __init__(self, inducing=None, num_inducing=100, bin_along=None, bin_count=10, seed=12345):
        if inducing is not None:
            inducing = _as_matrix(inducing)
            num_inducing = inducing.shape[0]
        self.inducing = inducing
        self.num_inducing = int(num_inducing)
        self.bin_along = bin_along
        self.bin_count = min(int(bin_count), self.num_inducing)
        self.seed = seed


# ---------------------------------------------------------------------------
# the GP object (R/gp.R)
# ---------------------------------------------------------------------------

class _Fit:
    """Device-resident fit (gp$fit, R/laplace.R:119): freed with the handle."""

    def __init__(self, handle, n, log_evidence, iters, approx):
        self.handle = handle
        self.n = n
        self.log_evidence = log_evidence
        self.iters = iters
        self.type = "analytic"
        self.approx = approx
        self._fin = weakref.finalize(self, _free_model, handle.value)

    def vector(self, which):
        out = np.empty(self.n)
        L.check(L.lib().gpl_model_vector(self.handle, which, L.dptr(out)))
        return out

    @property
    def fmean(self):
        return self.vector(L.VEC_FMEAN)

    @property
    def alpha(self):
        return self.vector(L.VEC_ALPHA)

    @property
    def z(self):
        return self.vector(L.VEC_Z)

    @property
    def V(self):
        return self.vector(L.VEC_V)

    @property
    def fvar(self):
        return self.vector(L.VEC_FVAR)

    @property
    def precision_info(self):
        """The precision the fit's results come from ("mixed" or "fp64"), whether a
        mixed fit was redone in FP64, and the last refinement's sweeps and relative
        correction."""
        p, f, k, r = C_int32(), C_int32(), C_int32(), C_double()
        L.check(L.lib().gpl_model_precision(self.handle, C_byref(p), C_byref(f), C_byref(k), C_byref(r)))
        return {"precision": "mixed" if p.value == L.PREC_MIXED else "fp64", "refit_fp64": bool(f.value),
                "refine_sweeps": int(k.value), "refine_rel": float(r.value)}

    @property
    def C_chol(self):
        out = np.empty((self.n, self.n), order="F")
        L.check(L.lib().gpl_model_chol(self.handle, out.ctypes.data_as(L._dp)))
        return out


def _free_model(ptr):
    if ptr and L._lib is not None:
        L._lib.gpl_model_free(ptr)


def _free_fitc(ptr):
    if ptr and L._lib is not None:
        L._lib.gpl_fitc_free(ptr)


class _FitcFit:
    """Device-resident FITC fit (gp$fit of R/laplace.R:177-180 and
    R/ep.R:271-278: fmean, z, V, alpha, fvar and cavity moments for EP, the FITC
    diagonal and the log evidence)."""

    def __init__(self, handle, n, m, log_evidence, iters, approx):
        self.handle = handle
        self.n, self.m = n, m
        self.log_evidence = log_evidence
        self.iters = iters
        self.type = "analytic"
        self.approx = approx
        self._fin = weakref.finalize(self, _free_fitc, handle.value if handle is not None else None)

    def vector(self, which):
        if self.handle is None:
            raise RuntimeError("the FITC fit failed (log_evidence = -Inf); there is nothing to read")
        out = np.empty(self.n)
        L.check(L.lib().gpl_fitc_vector(self.handle, which, L.dptr(out)))
        return out

    fmean = property(lambda self: self.vector(L.VEC_FMEAN))
    alpha = property(lambda self: self.vector(L.VEC_ALPHA))
    z = property(lambda self: self.vector(L.VEC_Z))
    V = property(lambda self: self.vector(L.VEC_V))
    fvar = property(lambda self: self.vector(L.VEC_FVAR))
    cavity_mean = property(lambda self: self.vector(L.VEC_CAVITY_MEAN))
    cavity_var = property(lambda self: self.vector(L.VEC_CAVITY_VAR))
    diag = property(lambda self: self.vector(7))


class GP:
    def __init__(self, cfs, lik, method, approx):
        self.cfs = cfs
        self.lik = lik
        self.method = method
        self.approx = approx
        self.fitted = False
        self.fit = None
        self.x = None
        self.precision = "fp64"

    def copy(self):
        g = GP(copy.deepcopy(self.cfs), copy.deepcopy(self.lik), self.method, self.approx)
        g.fitted, g.fit, g.x, g.precision = self.fitted, self.fit, self.x, self.precision
        return g


def gp_init(cfs=None, lik=None, method=None, approx=None):
    """R/gp.R:52-71."""
    cfs = cf_sexp() if cfs is None else cfs
    lik = lik_gaussian() if lik is None else lik
    method = method_full() if method is None else method
    approx = approx_laplace() if approx is None else approx
    if isinstance(lik, lik_gaussian) and isinstance(approx, approx_ep):
        raise ValueError("approx_ep disabled for lik_gaussian: use approx_laplace instead, because it gives exact "
                         "results")
    if not isinstance(method, (method_full, method_fitc)):
        raise NotImplementedError("method_full and method_fitc are provided")
    if isinstance(cfs, Cf):
        cfs = [cfs]
    return GP(list(cfs), lik, method, approx)


# ---- parameters (R/cf.R:327-382, R/gp.R:150-165) -------------------------

def _slots(gp, free=True):
    out = []
    for cf in gp.cfs:
        out.extend(cf.slots() if free else cf.all_slots())
    out.extend(gp.lik.slots() if free else gp.lik.all_slots())
    return out


def _slot_label(obj, nm, i, owner_label):
    return f"{owner_label}.{nm}" if i is None else f"{owner_label}.{nm}[{i}]"


def _labels(gp):
    labs = []

    def visit(cf):
        if isinstance(cf, cf_prod):
            for c in cf.cfs:
                visit(c)
            return
        lab = cf.label()
        for (o, nm, i) in cf.slots():
            labs.append(_slot_label(o, nm, i, lab))

    for cf in gp.cfs:
        visit(cf)
    for (o, nm, i) in gp.lik.slots():
        labs.append(_slot_label(o, nm, i, gp.lik.label()))
    return labs


def get_param(gp):
    """log of every free hyperparameter, as an ordered {name: value} dict."""
    vals = []
    for (o, nm, i) in _slots(gp):
        v = getattr(o, nm)
        vals.append(math.log(float(v if i is None else np.atleast_1d(v)[i])))
    return dict(zip(_labels(gp), vals))


def set_param(gp, param):
    """Returns a copy with exp(param) set and the fit invalidated (R/gp.R:158-165)."""
    g = gp.copy()
    vals = list(param.values()) if isinstance(param, dict) else list(np.asarray(param, dtype=float))
    sl = _slots(g)
    if len(vals) != len(sl):
        raise ValueError(f"expected {len(sl)} parameters, got {len(vals)}")
    for (o, nm, i), v in zip(sl, vals):
        if i is None:
            setattr(o, nm, math.exp(v))
        else:
            arr = np.array(np.atleast_1d(getattr(o, nm)), dtype=float)
            arr[i] = math.exp(v)
            setattr(o, nm, arr)
    g.fitted = False
    return g


def lpdf_prior(gp):
    """R/gp.R:167-169."""
    lp = 0.0
    for (o, nm, i) in _slots(gp):
        v = getattr(o, nm)
        p = math.log(float(v if i is None else np.atleast_1d(v)[i]))
        lp += o.priors.get(nm, prior_logunif()).lpdf(p)
    return lp


def _lpdf_prior_grad(gp):
    out = []
    for (o, nm, i) in _slots(gp):
        v = getattr(o, nm)
        p = math.log(float(v if i is None else np.atleast_1d(v)[i]))
        out.append(o.priors.get(nm, prior_logunif()).dlpdf(p))
    return np.array(out)


def is_fitted(gp):
    """R/gp.R:191-205."""
    if gp.fit is not None:
        if not gp.fitted:
            raise RuntimeError("The GP object seems to contain a posterior fit, but is not refitted after setting "
                               "new hyperparameter values. Please refit using gp_fit after calling set_param.")
        return True
    return False


# ---- covariance program for the device ----------------------------------

_KIND = {"const": L.CF_CONST, "lin": L.CF_LIN, "sexp": L.CF_SEXP, "matern32": L.CF_MATERN32,
         "matern52": L.CF_MATERN52, "nn": L.CF_NN, "periodic": L.CF_PERIODIC}


def _prod_factors(cf):
    """The factors of a product, nested products expanded in order (a product of
    products is one product: .cf_prod, R/cf.R:491-497)."""
    out = []
    for c in cf.cfs:
        out.extend(_prod_factors(c) if isinstance(c, cf_prod) else [c])
    return out


def _flatten(cfs):
    """[(cf, op)] with products flattened; op = SUM starts a summand."""
    out = []
    for cf in cfs:
        if isinstance(cf, cf_prod):
            for k, c in enumerate(_prod_factors(cf)):
                out.append((c, L.OP_SUM if k == 0 else L.OP_PROD))
        else:
            out.append((cf, L.OP_SUM))
    return out


def _program(cfs, x):
    """Build the gpl_cf_term array and the prepared input matrix (column-major)."""
    terms = _flatten(cfs)
    if len(terms) > L.MAX_TERMS:
        raise ValueError("at most 8 covariance terms")
    arr = (L.CfTerm * len(terms))()
    blocks = []
    col = 0
    free_mask = []
    for k, (cf, op) in enumerate(terms):
        t = arr[k]
        t.kind = _KIND[cf.kind]
        t.op = op
        xc = cf.prepare(x)
        if cf.kind == "const":
            t.ncols = 0
            t.col0 = 0
        else:
            t.col0 = col
            t.ncols = xc.shape[1]
            if t.ncols > L.MAX_COLS:
                raise ValueError("a covariance term reads at most 16 input columns")
            blocks.append(xc)
            col += xc.shape[1]
        if cf.kind in ("sexp", "matern32", "matern52"):
            ls = np.atleast_1d(cf.lscale)
            t.n_lscale = ls.size
            t.lscale[:ls.size] = ls
            t.magn = cf.magn
        elif cf.kind == "periodic":
            b = cf.base
            t.base_kind = _KIND[b.kind]
            ls = np.atleast_1d(b.lscale)
            t.n_lscale = ls.size
            t.lscale[:ls.size] = ls
            t.magn = b.magn
            t.period = cf.period
        elif cf.kind == "nn":
            t.sigma0, t.sigma, t.magn = cf.sigma0, cf.sigma, cf.magn
        else:
            t.magn = cf.magn
        free_mask.extend(0 if _fixed(o.priors.get(nm)) else 1 for (o, nm, _i) in cf.all_slots())
    xp = np.hstack(blocks) if blocks else np.zeros((_as_matrix(x).shape[0], 0))
    return arr, len(terms), np.asfortranarray(xp), free_mask


def _cfind_list(cfs, cfind):
    """0-based indices of top-level covariance functions (R's cfind, 1-based there)."""
    sel = [int(i) for i in np.atleast_1d(np.asarray(cfind, dtype=int))]
    if not sel or min(sel) < 0 or max(sel) >= len(cfs):
        raise IndexError(f"cfind must index the {len(cfs)} covariance functions (0-based)")
    return sel


def _cfind_terms(cfs, cfind):
    """The device term indices (products flattened, as _flatten orders them) of cfind."""
    groups, k = [], 0
    for cf in cfs:
        nt = len(_prod_factors(cf)) if isinstance(cf, cf_prod) else 1
        groups.append(list(range(k, k + nt)))
        k += nt
    idx = [t for i in _cfind_list(cfs, cfind) for t in groups[i]]
    if len(idx) > L.MAX_TERMS:
        raise ValueError("at most 8 covariance terms")
    return (C_int32 * len(idx))(*idx), len(idx)


def eval_cf(cfs, x1, x2, diag_only=False, cfind=None):
    """K between two input sets on the device (R/cf.R:402-411); cfind keeps the
    listed covariance functions (0-based)."""
    if isinstance(cfs, Cf):
        cfs = [cfs]
    if cfind is not None:
        cfs = [cfs[i] for i in _cfind_list(cfs, cfind)]
    t1, nt, x1p, _ = _program(cfs, x1)
    _, _, x2p, _ = _program(cfs, x2)
    n1, n2 = x1p.shape[0], x2p.shape[0]
    out = np.empty(min(n1, n2)) if diag_only else np.empty((n1, n2), order="F")
    L.check(L.lib().gpl_eval_cf(L.ctx(), t1, nt, L.dptr(x1p), n1, L.dptr(x2p), n2, x1p.shape[1], int(diag_only),
                                out.ctypes.data_as(L._dp)))
    return out


# ---- gp_fit / gp_energy / gp_pred ---------------------------------------

_LIKMAP = {"gaussian": L.LIK_GAUSSIAN, "bernoulli": L.LIK_BERNOULLI, "binomial": L.LIK_BINOMIAL,
           "poisson": L.LIK_POISSON, "betabinom": L.LIK_BETABINOM}
_LINKMAP = {"identity": L.LINK_IDENTITY, "logit": L.LINK_LOGIT, "probit": L.LINK_PROBIT, "log": L.LINK_LOG}


def _lik_struct(lik):
    if lik.kind not in _LIKMAP:
        raise NotImplementedError(f"lik_{lik.kind} is not on the device path")
    s = L.Lik()
    s.kind = _LIKMAP[lik.kind]
    s.link = _LINKMAP[lik.link]
    s.sigma = getattr(lik, "sigma", 0.0)
    s.phi = getattr(lik, "phi", 0.0)
    return s


def _scale_cols(x):
    """R's scale(): centre by the column means, divide by the column sd (n - 1)."""
    mu = x.mean(axis=0)
    sd = x.std(axis=0, ddof=1) if x.shape[0] > 1 else np.ones(x.shape[1])
    sd = np.where(sd > 0, sd, 1.0)
    return (x - mu) / sd, mu, sd


def _kmeans_centres(xs, k, rng, iters=100):
    """k-means centres of the rows of xs: k-means++ seeding, then Lloyd
    iterations until the assignment is stable.  The reference calls
    stats::kmeans (Hartigan-Wong) with starts from R's generator, so the
    centres are not the reference's; FITC given the same inducing inputs is."""
    n = xs.shape[0]
    if k > n:
        raise ValueError("more cluster centers than distinct data points")
    sq = np.sum(xs * xs, axis=1)
    first = int(rng.integers(n))
    C = [xs[first]]
    d2 = np.sum((xs - xs[first]) ** 2, axis=1)
    for _ in range(1, k):
        tot = d2.sum()
        i = int(rng.choice(n, p=d2 / tot)) if tot > 0 else int(rng.integers(n))
        C.append(xs[i])
        d2 = np.minimum(d2, np.sum((xs - xs[i]) ** 2, axis=1))
    C = np.array(C)
    lab = None
    for _ in range(iters):
        new = np.empty(n, dtype=np.int64)
        for a in range(0, n, 8192):  # distances in row chunks
            b = min(n, a + 8192)
            d = sq[a:b, None] - 2.0 * xs[a:b] @ C.T + np.sum(C * C, axis=1)[None, :]
            new[a:b] = np.argmin(d, axis=1)
        if lab is not None and np.array_equal(new, lab):
            break
        lab = new
        cnt = np.bincount(lab, minlength=k)
        for j in range(xs.shape[1]):
            s_ = np.bincount(lab, weights=xs[:, j], minlength=k)
            C[cnt > 0, j] = s_[cnt > 0] / cnt[cnt > 0]
    return C


def get_inducing(gp, x):
    """Inducing inputs of a method_fitc model (R/gp_fit.R:195-240): the given
    ones, or k-means centres of the scaled inputs, per bin of input column
    bin_along (0-based) when it is set."""
    meth = gp.method
    if meth.inducing is not None:
        return _as_matrix(meth.inducing)
    x = _as_matrix(x)
    rng = np.random.default_rng(meth.seed)
    k = meth.num_inducing
    if meth.bin_along is None:
        xs, mu, sd = _scale_cols(x)
        return _kmeans_centres(xs, k, rng) * sd + mu
    v = x[:, meth.bin_along]
    eps = 2 * np.finfo(float).eps
    cuts = np.linspace(v.min() - eps, v.max() + eps, meth.bin_count + 1)  # R/gp_fit.R:243-259
    bins = [x[(v > cuts[i]) & (v < cuts[i + 1])] for i in range(meth.bin_count)]
    order = np.argsort([-len(b) for b in bins], kind="stable")
    bins = [bins[i] for i in order]
    per = np.full(meth.bin_count, k // meth.bin_count)
    per[: k - per.sum()] += 1
    out = []
    for b, kb in zip(bins, per):
        bs, mu, sd = _scale_cols(b)
        out.append(_kmeans_centres(bs, int(kb), rng) * sd + mu)
    return np.vstack(out)


def _fit_fitc(g, x, y, terms, nt, xp, tr, off, lk, o):
    """fit_laplace.method_fitc (R/gp_fit.R:101-123) through gpl_fitc_fit."""
    if g.precision == "mixed":
        raise NotImplementedError("FITC runs in FP64")
    g.method = copy.copy(g.method)
    z = get_inducing(g, x)
    g.method.inducing = z
    _, _, zp, _ = _program(g.cfs, z)
    n, m = y.shape[0], z.shape[0]
    handle = L._vp()
    logz = C_double()
    rc = L.lib().gpl_fitc_fit(L.ctx(), terms, nt, L.dptr(xp), n, xp.shape[1], L.dptr(zp), m, L.dptr(y), L.dptr(tr),
                              L.dptr(off), C_byref(lk), C_byref(o), C_byref(handle), C_byref(logz))
    if rc in (L.GPL_ERR_NOT_PD, L.GPL_ERR_NAN):
        # the reference catches the error, prints it and keeps log_evidence = -Inf (R/gp_fit.R:113-121)
        print(L.lib().gpl_last_error().decode())
        fit = _FitcFit(None, n, m, -math.inf, 0, o.approx)
    else:
        if rc not in (L.GPL_OK, L.GPL_ERR_NOCONV):
            L.check(rc)
        iters = C_int32()
        L.lib().gpl_fitc_info(handle, None, None, C_byref(iters), None)
        fit = _FitcFit(handle, n, m, float(logz.value), int(iters.value), o.approx)
        if rc == L.GPL_ERR_NOCONV:
            warnings.warn(L.lib().gpl_last_error().decode())
    g.fit = fit
    g.x = x
    g.fitted = True
    g._trials, g._offset = tr, off
    return g


def gp_fit(gp, x, y, trials=None, offset=None, jitter=None, precision=None, want_grad=False, refine_iters=None):
    """R/gp_fit.R:54-60: learn input scales, build K + jitter I on the device and
    run the Laplace Newton loop or EP; returns a new fitted GP.  want_grad forms
    L^-1 during the fit (the gradient needs it anyway).  refine_iters caps the
    mixed form's refinement sweeps per solve (default 8); a solve that does not
    converge within them makes the device redo the fit in FP64."""
    g = gp.copy()
    x = _as_matrix(x)
    y = np.ascontiguousarray(np.asarray(y, dtype=np.float64).ravel())
    if x.shape[0] != y.shape[0]:
        raise ValueError("x and y have different numbers of observations")
    for cf in g.cfs:
        cf.learn_scales(x)
    terms, nt, xp, _ = _program(g.cfs, x)
    if isinstance(g.lik, lik_binomial) and trials is None:
        raise ValueError("trials must be provided for the binomial likelihood.")
    if isinstance(g.lik, lik_betabinom) and trials is None:
        raise ValueError("trials must be provided for the beta binomial likelihood.")
    if isinstance(g.lik, lik_bernoulli) and not isinstance(g.lik, lik_binomial):
        trials = None
    tr = None if trials is None else np.ascontiguousarray(np.asarray(trials, dtype=np.float64).ravel())
    off = None if offset is None else np.ascontiguousarray(np.broadcast_to(np.asarray(offset, float), y.shape))
    o = L.Opts()
    L.lib().gpl_opts_default(C_byref(o))
    if isinstance(g.approx, approx_ep):
        o.approx = L.APPROX_EP
        o.maxiter = g.approx.maxiter
        o.damping = g.approx.damping
        o.quad_order = g.approx.quad_order
    else:
        o.approx = L.APPROX_LAPLACE
        o.maxiter = g.approx.maxiter
        o.tol = g.approx.tol
    o.jitter = 1e-6 if jitter is None else float(jitter)
    o.flags = (1 if want_grad else 0) | (2 if getattr(g.approx, "analytic_w", False) else 0)
    if refine_iters is not None:
        o.refine_iters = int(refine_iters)
    prec = precision or g.precision
    o.precision = L.PREC_MIXED if prec == "mixed" else L.PREC_FP64
    g.precision = prec
    lk = _lik_struct(g.lik)
    if isinstance(g.method, method_fitc):
        return _fit_fitc(g, x, y, terms, nt, xp, tr, off, lk, o)
    handle = L._vp()
    logz = C_double()
    n = y.shape[0]
    rc = L.lib().gpl_fit(L.ctx(), terms, nt, L.dptr(xp), n, xp.shape[1], L.dptr(y), L.dptr(tr), L.dptr(off),
                         C_byref(lk), C_byref(o), C_byref(handle), C_byref(logz))
    if rc != L.GPL_OK and not handle.value:
        L.check(rc)
    iters = C_int32()
    L.lib().gpl_model_info(handle, None, C_byref(iters), None)
    fit = _Fit(handle, n, float(logz.value), int(iters.value), o.approx)
    if rc == L.GPL_ERR_NOCONV:
        warnings.warn(L.lib().gpl_last_error().decode())
    elif rc == L.GPL_ERR_NAN:
        fit.log_evidence = -math.inf
    elif rc != L.GPL_OK:
        L.check(rc)
    g.fit = fit
    g.x = x
    g.fitted = True
    g._trials, g._offset = tr, off
    return g


def gp_energy(gp, include_prior=True):
    """-log evidence (- log prior): R/gp.R:138-147."""
    if not is_fitted(gp):
        raise RuntimeError("The GP must be fitted. Call gp_fit first.")
    e = -gp.fit.log_evidence
    if include_prior:
        e -= lpdf_prior(gp)
    return e


def gp_energy_grad(gp, include_prior=True):
    """d gp_energy / d get_param(gp), computed on the device: Gaussian likelihood
    (any precision); EP and Laplace with approx_laplace(analytic_w=True) in FP64,
    whose likelihood parameters (the beta-binomial phi) must be fixed."""
    if not is_fitted(gp):
        raise RuntimeError("The GP must be fitted. Call gp_fit first.")
    if isinstance(gp.method, method_fitc):
        raise NotImplementedError("the energy gradient is provided for method_full")
    terms, nt, _, mask = _program(gp.cfs, gp.x)
    ng = L.lib().gpl_grad_len(terms, nt, _LIKMAP[gp.lik.kind])
    latent = gp.lik.kind != "gaussian"
    if latent and any(not _fixed(gp.lik.priors.get(nm)) for nm in gp.lik.param_names):
        raise NotImplementedError("the gradient in the likelihood's parameters is provided for lik_gaussian only; "
                                  "fix them with prior_fixed()")
    full = np.empty(ng)
    L.check(L.lib().gpl_energy_grad(gp.fit.handle, L.dptr(full)))
    if not latent:
        mask = mask + [0 if _fixed(gp.lik.priors.get(nm)) else 1 for nm in gp.lik.param_names]
    g = full[np.asarray(mask, dtype=bool)]
    if include_prior:
        g = g - _lpdf_prior_grad(gp)
    return g


def gp_pred(gp, xnew, var=False, quantiles=None, transform=False, cfind=None, jitter=None, quad_order=15,
            offset=None):
    """R/gp_pred.R:88-122 for method_full: posterior if fitted, prior otherwise.
    cfind (0-based) keeps only those covariance functions in K_t and K_tt."""
    if quantiles is not None or transform:
        var = True
    xt = _as_matrix(xnew)
    m = xt.shape[0]
    fitc = isinstance(gp.method, method_fitc)
    if fitc and cfind is not None:
        raise NotImplementedError("cfind with method_fitc is not on the device path")
    if not is_fitted(gp):  # R/gp_pred.R:142-184
        if fitc and var and gp.method.inducing is None:
            raise ValueError("Inducing points not set yet.")
        mean = np.zeros(m)
        v = eval_cf(gp.cfs, xt, xt, diag_only=True, cfind=cfind) if var else None
    elif fitc:  # R/gp_pred.R:234-290
        if gp.fit.handle is None:
            raise RuntimeError("the FITC fit failed (log_evidence = -Inf)")
        _, _, xtp, _ = _program(gp.cfs, xt)
        mean = np.empty(m)
        v = np.empty(m) if var else None
        L.check(L.lib().gpl_fitc_pred(gp.fit.handle, L.dptr(xtp), m, int(bool(var)), L.dptr(mean), L.dptr(v)))
    else:
        _, _, xtp, _ = _program(gp.cfs, xt)
        mean = np.empty(m)
        v = np.empty(m) if var else None
        if cfind is None:
            L.check(L.lib().gpl_pred(gp.fit.handle, L.dptr(xtp), m, int(bool(var)), L.dptr(mean), L.dptr(v)))
        else:
            sel, ns = _cfind_terms(gp.cfs, cfind)
            L.check(L.lib().gpl_pred_terms(gp.fit.handle, sel, ns, L.dptr(xtp), m, int(bool(var)), L.dptr(mean),
                                           L.dptr(v)))
    if offset is not None:
        mean = mean + offset
    out = {"mean": mean}
    if var:
        out["var"] = v
    if quantiles is not None:
        from scipy.stats import norm
        q = np.column_stack([norm.ppf(qq, loc=mean, scale=np.sqrt(v)) for qq in np.atleast_1d(quantiles)])
        if transform:
            q = gp.lik.response(q)
        out["quantiles"] = q
    if transform:
        xg, wg = _gh(quad_order)
        fg = mean[:, None] + np.sqrt(v)[:, None] * xg[None, :]
        ft = gp.lik.response(fg)
        mt = ft @ wg
        out["mean"] = mt
        out["var"] = ((ft - mt[:, None]) ** 2) @ wg
    return out


def gp_draw(gp, xnew, draws=1, transform=True, target=False, marginal=False, cfind=None, jitter=None, seed=None,
            trials=None, offset=None):
    """Posterior draws at xnew: R/gp_draw.R:42-140 for method_full and
    method_fitc, returned as a (len(xnew), draws) array.

    marginal=False draws mean + chol(cov + jitter I) z on the device
    (gpl_draw: full predictive covariance, its Cholesky factor and the triangular
    product).  marginal=True uses the predictive variances only.  The standard
    normal draws z come from numpy's Generator(PCG64(seed)); R's rnorm stream
    cannot be reproduced.  transform maps to the response scale; target draws
    observations (R/lik.R:502-537).
    """
    fitted = is_fitted(gp)
    fitc = isinstance(gp.method, method_fitc)
    if fitc and cfind is not None:
        raise NotImplementedError("cfind with method_fitc is not on the device path")
    xt = _as_matrix(xnew)
    m = xt.shape[0]
    jitter = 1e-6 if jitter is None else float(jitter)
    rng = np.random.default_rng(seed)
    z = np.asfortranarray(rng.standard_normal((m, draws)))
    if not fitted:  # gp_draw_prior (R/gp_draw.R:56-67)
        if fitc:
            if gp.method.inducing is None:
                raise ValueError("Inducing points not set yet.")
            raise NotImplementedError("prior draws for method_fitc are not on the device path")
        cfs = gp.cfs if cfind is None else [gp.cfs[i] for i in _cfind_list(gp.cfs, cfind)]
        terms, nt, xtp, _ = _program(cfs, xt)
        f = np.empty((m, draws), order="F")
        L.check(L.lib().gpl_draw_prior(L.ctx(), terms, nt, L.dptr(xtp), m, xtp.shape[1], jitter,
                                       z.ctypes.data_as(L._dp), draws, f.ctypes.data_as(L._dp)))
    elif marginal:
        pr = gp_pred(gp, xt, var=True, cfind=cfind)
        f = pr["mean"][:, None] + np.sqrt(pr["var"])[:, None] * z
    else:
        _, _, xtp, _ = _program(gp.cfs, xt)
        f = np.empty((m, draws), order="F")
        if fitc:  # R/gp_draw.R:126-140
            if gp.fit.handle is None:
                raise RuntimeError("the FITC fit failed (log_evidence = -Inf)")
            L.check(L.lib().gpl_fitc_draw(gp.fit.handle, L.dptr(xtp), m, jitter, z.ctypes.data_as(L._dp), draws,
                                          f.ctypes.data_as(L._dp)))
        elif cfind is None:
            L.check(L.lib().gpl_draw(gp.fit.handle, L.dptr(xtp), m, jitter, z.ctypes.data_as(L._dp), draws,
                                     f.ctypes.data_as(L._dp)))
        else:
            sel, ns = _cfind_terms(gp.cfs, cfind)
            L.check(L.lib().gpl_draw_terms(gp.fit.handle, sel, ns, L.dptr(xtp), m, jitter, z.ctypes.data_as(L._dp),
                                           draws, f.ctypes.data_as(L._dp)))
    if offset is not None:
        f = f + np.asarray(offset, dtype=float)[:, None]
    if target:
        mu = gp.lik.response(f)
        if isinstance(gp.lik, lik_gaussian):
            return f + gp.lik.sigma * rng.standard_normal(f.shape)
        if isinstance(gp.lik, lik_poisson):
            return rng.poisson(mu).astype(float)
        t = np.ones(m) if trials is None else np.asarray(trials, dtype=float)
        if isinstance(gp.lik, lik_betabinom):  # generate_target.lik_betabinom, R/lik.R:526-533
            mu = rng.beta(mu / gp.lik.phi, (1 - mu) / gp.lik.phi)
        return rng.binomial(t.astype(np.int64)[:, None], mu).astype(float)
    if transform:
        return gp.lik.response(f)
    return f


# ---------------------------------------------------------------------------
# gp_save / gp_load (R/saveload.R:32-41)
# ---------------------------------------------------------------------------

_FILE_FORMAT = "gplite_amd-gp-1"


def _saveable():
    return {c.__name__: c for c in (cf_const, cf_lin, cf_sexp, cf_matern32, cf_matern52, cf_nn, cf_periodic, cf_prod,
                                    lik_gaussian, lik_bernoulli, lik_binomial, lik_betabinom, lik_poisson, prior_fixed,
                                    prior_logunif, prior_lognormal, prior_half_t, approx_laplace, approx_ep,
                                    method_full, method_fitc)}


def _encode(obj, classes):
    if obj is None or isinstance(obj, (bool, int, float, str)):
        return obj
    if isinstance(obj, np.ndarray):
        return {"__ndarray__": obj.tolist(), "dtype": str(obj.dtype)}
    if isinstance(obj, np.generic):
        return obj.item()
    if isinstance(obj, (list, tuple)):
        return [_encode(v, classes) for v in obj]
    if isinstance(obj, dict):
        return {"__map__": {k: _encode(v, classes) for k, v in obj.items()}}
    name = type(obj).__name__
    if classes.get(name) is type(obj):
        return {"__class__": name, "state": {k: _encode(v, classes) for k, v in vars(obj).items()}}
    raise TypeError(f"gp_save: cannot store an object of type {name}")


def _decode(obj, classes):
    if isinstance(obj, list):
        return [_decode(v, classes) for v in obj]
    if not isinstance(obj, dict):
        return obj
    if "__ndarray__" in obj:
        return np.asarray(obj["__ndarray__"], dtype=np.dtype(obj["dtype"]))
    if "__map__" in obj:
        return {k: _decode(v, classes) for k, v in obj["__map__"].items()}
    cls = classes.get(obj.get("__class__"))
    if cls is None:
        raise ValueError(f"gp_load: unknown object class {obj.get('__class__')!r}")
    inst = cls.__new__(cls)  # state only: no constructor, no code from the file
    inst.__dict__.update({k: _decode(v, classes) for k, v in obj["state"].items()})
    return inst


def gp_save(gp, filename):
    """R/saveload.R:32-34: store the model, fitted or not, in one file.

    The model structure (covariance functions with their learned normalisation,
    likelihood, priors, approximation) is stored as JSON and the numbers as
    arrays of an .npz archive; nothing is pickled.  A fitted model also stores
    the gp$fit fields of R/laplace.R:119 and R/ep.R:140-145 (C_chol, fmean,
    alpha, z, V, fvar, cavity moments, log_evidence), read back from the device,
    and its training inputs.  A fit made stale by set_param is not stored."""
    import json
    classes = _saveable()
    head = {"format": _FILE_FORMAT, "cfs": _encode(gp.cfs, classes), "lik": _encode(gp.lik, classes),
            "method": _encode(gp.method, classes), "approx": _encode(gp.approx, classes),
            "precision": gp.precision, "fitted": bool(gp.fit is not None and gp.fitted)}
    arrays = {}
    if head["fitted"] and isinstance(gp.method, method_fitc):
        raise NotImplementedError("saving a fitted method_fitc model is not provided")
    if head["fitted"]:
        f = gp.fit
        head.update(n=int(f.n), log_evidence=float(f.log_evidence), iters=int(f.iters))
        arrays["x"] = np.asarray(gp.x, dtype=np.float64)
        arrays["C_chol"] = f.C_chol
        for nm, w in (("fmean", L.VEC_FMEAN), ("alpha", L.VEC_ALPHA), ("z", L.VEC_Z), ("V", L.VEC_V),
                      ("fvar", L.VEC_FVAR), ("cavity_mean", L.VEC_CAVITY_MEAN),
                      ("cavity_var", L.VEC_CAVITY_VAR)):
            arrays[nm] = f.vector(w)
        for nm in ("_trials", "_offset"):
            v = getattr(gp, nm, None)
            if v is not None:
                arrays[nm.lstrip("_")] = np.asarray(v, dtype=np.float64)
    arrays["header"] = np.array(json.dumps(head))
    with open(filename, "wb") as fh:
        np.savez(fh, **arrays)


def gp_load(filename):
    """R/saveload.R:37-41: the model stored by gp_save.  A fitted model is
    restored on the device from its saved fields (gpl_model_restore: no refit),
    in FP64 whatever precision it was fitted in."""
    import json
    classes = _saveable()
    with np.load(filename, allow_pickle=False) as z:
        head = json.loads(str(z["header"]))
        if head.get("format") != _FILE_FORMAT:
            raise ValueError(f"gp_load: {filename} is not a saved gplite_amd model")
        arrays = {k: z[k] for k in z.files if k != "header"}
    g = GP(_decode(head["cfs"], classes), _decode(head["lik"], classes), _decode(head["method"], classes),
           _decode(head["approx"], classes))
    g.precision = head["precision"]
    if not head["fitted"]:
        return g
    x = _as_matrix(arrays["x"])
    n = int(head["n"])
    terms, nt, xp, _ = _program(g.cfs, x)
    o = L.Opts()
    L.lib().gpl_opts_default(C_byref(o))
    o.approx = L.APPROX_EP if isinstance(g.approx, approx_ep) else L.APPROX_LAPLACE
    o.flags = 2 if getattr(g.approx, "analytic_w", False) else 0
    lk = _lik_struct(g.lik)
    vecs = [np.ascontiguousarray(arrays[nm], dtype=np.float64) for nm in
            ("fmean", "alpha", "z", "V", "fvar", "cavity_mean", "cavity_var")]
    vptr = (L._dp * 7)(*[L.dptr(v) for v in vecs])
    cc = np.asfortranarray(arrays["C_chol"], dtype=np.float64)
    if cc.shape != (n, n) or x.shape[0] != n or any(v.shape != (n,) for v in vecs):
        raise ValueError("gp_load: inconsistent sizes in the saved fit")
    handle = L._vp()
    L.check(L.lib().gpl_model_restore(L.ctx(), terms, nt, L.dptr(xp), n, xp.shape[1], C_byref(lk), C_byref(o),
                                      cc.ctypes.data_as(L._dp), vptr, float(head["log_evidence"]),
                                      int(head["iters"]), C_byref(handle)))
    g.fit = _Fit(handle, n, float(head["log_evidence"]), int(head["iters"]), o.approx)
    g.x = x
    g.fitted = True
    g._trials = arrays.get("trials")
    g._offset = arrays.get("offset")
    return g


def _loglik_points(lik, f, y, trials=None, offset=None):
    """Pointwise observation log density log p(y_i | f_ik) (R/lik.R:297-366);
    f has shape (n, q) and broadcasts against y[:, None]."""
    from scipy.special import gammaln, ndtr
    f = np.asarray(f, dtype=float)
    if offset is not None:
        f = f + np.asarray(offset, dtype=float)[:, None]
    yy = np.asarray(y, dtype=float)[:, None]
    if isinstance(lik, lik_gaussian):
        s = lik.sigma
        return -0.5 * math.log(2 * math.pi) - math.log(s) - 0.5 * ((yy - f) / s) ** 2
    if isinstance(lik, lik_bernoulli):
        t = np.ones_like(yy) if (trials is None or not isinstance(lik, lik_binomial)) else \
            np.asarray(trials, dtype=float)[:, None]
        mu = ndtr(f) if lik.link == "probit" else 1.0 / (1.0 + np.exp(-f))
        with np.errstate(divide="ignore", invalid="ignore"):
            lc = gammaln(t + 1) - gammaln(yy + 1) - gammaln(t - yy + 1)
            return lc + np.where(yy > 0, yy * np.log(mu), 0.0) + np.where(t - yy > 0, (t - yy) * np.log1p(-mu), 0.0)
    if isinstance(lik, lik_poisson):
        mu = np.exp(f)
        return np.where(yy > 0, yy * f, 0.0) - mu - gammaln(yy + 1)
    if isinstance(lik, lik_betabinom):  # R/lik.R:332-352
        t = np.asarray(trials, dtype=float)[:, None]
        mu = ndtr(f) if lik.link == "probit" else 1.0 / (1.0 + np.exp(-f))
        a, b = mu / lik.phi, (1 - mu) / lik.phi
        return (gammaln(t + 1) - gammaln(yy + 1) - gammaln(t - yy + 1) + gammaln(yy + a) + gammaln(t - yy + b)
                - gammaln(t + a + b) + gammaln(a + b) - gammaln(a) - gammaln(b))
    raise NotImplementedError(type(lik).__name__)


def gp_loo(gp, x, y, quadrature=True, quad_order=11, draws=4000, seed=None, trials=None, offset=None):
    """Approximate leave-one-out predictive log densities: R/gp_loo.R:73-110
    (method_full and method_fitc).  The LOO latent posteriors come from the
    fitted model: Laplace uses the device predictive variance at the training
    inputs (R/gp_loo.R:120-130; for FITC with the diagonal correction,
    train = TRUE); EP uses the cavity moments (:132-134).  The
    predictive density is integrated with Gauss-Hermite quadrature, or by
    Monte Carlo with `draws` samples."""
    if not is_fitted(gp):
        raise ValueError("The provied GP model is not fitted. Please call gp_fit (or gp_optim) first.")
    fitc = isinstance(gp.method, method_fitc)
    if fitc and gp.fit.handle is None:
        raise RuntimeError("the FITC fit failed (log_evidence = -Inf)")
    y = np.asarray(y, dtype=float).ravel()
    n = len(y)
    if isinstance(gp.approx, approx_ep):
        m = gp.fit.vector(L.VEC_CAVITY_MEAN)
        v = gp.fit.vector(L.VEC_CAVITY_VAR)
    else:
        fhat, z, V = gp.fit.fmean, gp.fit.z, gp.fit.V
        if fitc:  # gp_pred_post.method_fitc with train = TRUE (R/gp_pred.R:270-285)
            post_var = np.empty(n)
            L.check(L.lib().gpl_fitc_train_var(gp.fit.handle, L.dptr(post_var)))
        else:
            post_var = gp_pred(gp, x, var=True)["var"]
        v = 1.0 / (1.0 / post_var - 1.0 / V)
        m = fhat - v * (z - fhat) / V
    sd = np.sqrt(v)
    if quadrature:
        xg, wg = _gh(quad_order)
        ll = _loglik_points(gp.lik, m[:, None] + sd[:, None] * xg[None, :], y, trials, offset)
        c = ll.max(axis=1, keepdims=True)
        loos = c[:, 0] + np.log(np.sum(wg[None, :] * np.exp(ll - c), axis=1))
    else:
        rng = np.random.default_rng(seed)
        ll = _loglik_points(gp.lik, m[:, None] + sd[:, None] * rng.standard_normal((n, draws)), y, trials, offset)
        c = ll.max(axis=1, keepdims=True)
        loos = c[:, 0] + np.log(np.sum(np.exp(ll - c), axis=1)) - math.log(draws)
    return {"loo": float(np.sum(loos)), "sd": float(n * np.std(loos, ddof=1) / math.sqrt(n)), "loos": loos}


def gp_compare(*loos, ref=None):
    """R/gp_loo.R:147-175: compare gp_loo results.  Row i holds
    n * mean(loos_i - loos_ref) and n * sd(loos_i - loos_ref) / sqrt(n); ref
    (0-based) defaults to the model with the largest loo.  Returns a (k, 2)
    array with columns loo-diff and se."""
    for a in loos:
        if not (isinstance(a, dict) and "loo" in a and "loos" in a):
            raise TypeError(f"Expected objects of type 'loores', but found an object with type '{type(a).__name__}'")
    loo = np.array([a["loo"] for a in loos], dtype=float)
    if ref is None:
        ref = int(np.argmax(loo))
    base = np.asarray(loos[ref]["loos"], dtype=float)
    n = len(base)
    res = np.empty((len(loos), 2))
    for i, a in enumerate(loos):
        d = np.asarray(a["loos"], dtype=float) - base
        res[i, 0] = n * d.mean()
        res[i, 1] = n * (d.std(ddof=1) / math.sqrt(n)) if n > 1 else float("nan")
    return res


def _gh(q):
    x, w = np.polynomial.hermite_e.hermegauss(q)
    return x, w / w.sum()


def gp_optim(gp, x, y, tol=1e-4, tol_param=1e-1, maxiter=500, restarts=1, verbose=True, warnings_=True,
             method="Nelder-Mead", **fit_args):
    """R/gp_optim.R:63-125: minimise gp_energy over log-parameters.  Every
    evaluation is a device refit.  method="L-BFGS-B" uses the device gradient
    (Gaussian likelihood)."""
    from scipy.optimize import minimize
    names = list(get_param(gp).keys())
    p0 = np.array(list(get_param(gp).values()))
    n = len(np.asarray(y).ravel())
    state = {"it": 0}

    def energy(p):
        g2 = gp_fit(set_param(gp, p), x, y, **fit_args)
        e = gp_energy(g2)
        state["it"] += 1
        if verbose:
            print(" ".join(f"{v:8.2f}" for v in p), f"{e:10.2f}", f"{state['it']:9d}")
        return e / n, g2

    if method == "L-BFGS-B":
        fit_args = dict(fit_args, want_grad=True)

        def fg(p):
            e, g2 = energy(p)
            return e, gp_energy_grad(g2) / n
        res = minimize(fg, p0, jac=True, method="L-BFGS-B", options={"maxiter": maxiter, "ftol": tol})
    else:
        # stats::optim's Nelder-Mead as R/gp_optim.R:87-95 calls it: fnscale = n, stop
        # when the simplex's energies lie within reltol * (|f(p0)| + reltol) of each
        # other (no parameter tolerance), maxit counts energy evaluations, and the
        # first simplex steps every coordinate by 0.1 max|p0| (0.1 if p0 = 0).
        f0 = energy(p0)[0]
        step = 0.1 * float(np.max(np.abs(p0))) if np.any(p0 != 0) else 0.1
        simplex = np.vstack([p0] + [p0 + step * np.eye(len(p0))[k] for k in range(len(p0))])
        res = minimize(lambda p: energy(p)[0], p0, method="Nelder-Mead",
                       options={"maxfev": maxiter, "maxiter": 10 * maxiter + 1000, "fatol": tol * (abs(f0) + tol),
                                "xatol": np.inf, "initial_simplex": simplex})
    if res.status == 1 and warnings_:  # evaluation or iteration cap (R/gp_optim.R:97-99)
        warnings.warn("Maximum number of iterations reached, optimization may not have converged.")
    popt = res.x
    g = gp_fit(set_param(gp, popt), x, y, **fit_args)
    if warnings_:
        e_opt = gp_energy(g)
        for k in range(len(popt)):  # check_convergence, R/gp_optim.R:127-160
            d = np.zeros_like(popt); d[k] = tol_param
            e1 = gp_energy(gp_fit(set_param(gp, popt + d), x, y, **fit_args))
            e2 = gp_energy(gp_fit(set_param(gp, popt - d), x, y, **fit_args))
            if e_opt - min(e1, e2) > 1e-2:
                if restarts > 0:
                    return gp_optim(g, x, y, tol, tol_param, maxiter, restarts - 1, verbose, warnings_, method,
                                    **fit_args)
                warnings.warn(f"Not all hyperparameters have reached convergence within tolerance tol_param = "
                              f"{tol_param}. Try increasing the number of restarts, or reducing tol.")
                break
    _ = names
    return g


