"""ctypes binding of the C ABI in include/gplite_amd.h.

The shared library is loaded from this package directory only; there is no
CPU fallback: if libgplite_amd.so is missing or no GPU is visible, every entry
point raises.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgplite_amd.so")

GPL_OK, GPL_ERR_ARG, GPL_ERR_NOT_PD, GPL_ERR_OOM, GPL_ERR_NOCONV, GPL_ERR_HIP, GPL_ERR_NAN, GPL_ERR_UNSUPPORTED = range(8)
CF_CONST, CF_LIN, CF_SEXP, CF_MATERN32, CF_MATERN52, CF_NN, CF_PERIODIC = range(7)
OP_SUM, OP_PROD = 0, 1
LIK_GAUSSIAN, LIK_BERNOULLI, LIK_BINOMIAL, LIK_POISSON, LIK_BETABINOM = range(5)
LINK_IDENTITY, LINK_LOGIT, LINK_PROBIT, LINK_LOG = range(4)
APPROX_LAPLACE, APPROX_EP = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
VEC_FMEAN, VEC_ALPHA, VEC_Z, VEC_V, VEC_FVAR, VEC_CAVITY_MEAN, VEC_CAVITY_VAR = range(7)
MAX_TERMS, MAX_COLS = 8, 16


class CfTerm(C.Structure):
    _fields_ = [("kind", C.c_int32), ("op", C.c_int32), ("col0", C.c_int32), ("ncols", C.c_int32),
                ("base_kind", C.c_int32), ("n_lscale", C.c_int32), ("magn", C.c_double),
                ("sigma0", C.c_double), ("sigma", C.c_double), ("period", C.c_double),
                ("lscale", C.c_double * (2 * MAX_COLS))]


class Lik(C.Structure):
    _fields_ = [("kind", C.c_int32), ("link", C.c_int32), ("sigma", C.c_double), ("phi", C.c_double)]


class Opts(C.Structure):
    _fields_ = [("approx", C.c_int32), ("maxiter", C.c_int32), ("tol", C.c_double), ("damping", C.c_double),
                ("damping_min", C.c_double), ("quad_order", C.c_int32), ("precision", C.c_int32),
                ("jitter", C.c_double), ("refine_iters", C.c_int32), ("flags", C.c_int32)]


class GplError(RuntimeError):
    def __init__(self, code, msg, minor=0):
        super().__init__(f"[gpl status {code}] {msg}")
        self.code = code
        self.minor = minor


_lib = None
_ctx = None
_lock = threading.Lock()

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


def _sig(lib):
    i64, i32 = C.c_int64, C.c_int32
    sigs = {
        "gpl_ctx_create": (C.c_int, [i32, C.POINTER(_vp)]),
        "gpl_ctx_destroy": (None, [_vp]),
        "gpl_last_error": (C.c_char_p, []),
        "gpl_last_error_minor": (i64, []),
        "gpl_device_name": (C.c_int, [_vp, C.c_char_p, i32]),
        "gpl_profile_enable": (C.c_int, [_vp, i32]),
        "gpl_profile_reset": (C.c_int, [_vp]),
        "gpl_profile_read": (C.c_int, [_vp, i32, C.c_char_p, _dp, C.POINTER(i64), _dp, _dp]),
        "gpl_opts_default": (None, [C.POINTER(Opts)]),
        "gpl_grad_len": (i32, [C.POINTER(CfTerm), i32, i32]),
        "gpl_cf_sexp": (C.c_int, [_vp, _dp, i64, _dp, i64, i64, _dp, i64, C.c_double, i32, _dp]),
        "gpl_cf_matern32": (C.c_int, [_vp, _dp, i64, _dp, i64, i64, _dp, i64, C.c_double, i32, _dp]),
        "gpl_cf_matern52": (C.c_int, [_vp, _dp, i64, _dp, i64, i64, _dp, i64, C.c_double, i32, _dp]),
        "gpl_cf_nn": (C.c_int, [_vp, _dp, i64, _dp, i64, i64, C.c_double, C.c_double, C.c_double, i32, _dp]),
        "gpl_eval_cf": (C.c_int, [_vp, C.POINTER(CfTerm), i32, _dp, i64, _dp, i64, i64, i32, _dp]),
        "gpl_fit": (C.c_int, [_vp, C.POINTER(CfTerm), i32, _dp, i64, i64, _dp, _dp, _dp, C.POINTER(Lik),
                              C.POINTER(Opts), C.POINTER(_vp), _dp]),
        "gpl_energy_grad": (C.c_int, [_vp, _dp]),
        "gpl_pred": (C.c_int, [_vp, _dp, i64, i32, _dp, _dp]),
        "gpl_model_vector": (C.c_int, [_vp, i32, _dp]),
        "gpl_model_chol": (C.c_int, [_vp, _dp]),
        "gpl_model_info": (C.c_int, [_vp, C.POINTER(i64), C.POINTER(i32), _dp]),
        "gpl_model_precision": (C.c_int, [_vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), _dp]),
        "gpl_pred_noalloc": (C.c_int, [_vp, _dp, i64, i32, _dp, _dp]),
        "gpl_model_trim": (C.c_int, [_vp]),
        "gpl_model_free": (None, [_vp]),
        "gpl_batch_create": (C.c_int, [_vp, i32, C.POINTER(CfTerm), i32, C.POINTER(_dp), C.POINTER(i64),
                                       C.POINTER(i64), C.POINTER(_dp), _dp, C.POINTER(Opts), C.POINTER(_vp)]),
        "gpl_batch_set_params": (C.c_int, [_vp, C.POINTER(CfTerm), _dp]),
        "gpl_batch_energy_grad": (C.c_int, [_vp, _dp, _dp]),
        "gpl_batch_grad_len": (i32, [_vp]),
        "gpl_batch_info": (C.c_int, [_vp, C.POINTER(i32), _dp]),
        "gpl_batch_status": (C.c_int, [_vp, C.POINTER(i32), _dp]),
        "gpl_fit_batch": (C.c_int, [_vp, i32, C.POINTER(CfTerm), i32, C.POINTER(_dp), C.POINTER(i64),
                                    C.POINTER(i64), C.POINTER(_dp), _dp, C.POINTER(Opts), C.POINTER(_vp), _dp,
                                    C.POINTER(i32)]),
        "gpl_batch_tune": (C.c_int, [_vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), _dp,
                                     C.POINTER(i32)]),
        "gpl_batch_set_tuning": (C.c_int, [_vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
        "gpl_batch_free": (None, [_vp]),
        "gpl_server_create": (C.c_int, [i32, i32, i64, C.POINTER(_vp)]),
        "gpl_server_submit_fit": (C.c_int, [_vp, C.POINTER(CfTerm), i32, _dp, i64, i64, _dp, _dp, _dp,
                                            C.POINTER(Lik), C.POINTER(Opts), C.POINTER(i64), C.POINTER(i64)]),
        "gpl_server_submit_pred": (C.c_int, [_vp, i64, _dp, i64, i32, _dp, _dp, C.POINTER(i64)]),
        "gpl_server_wait": (C.c_int, [_vp, i64, _dp, _dp, _dp, C.POINTER(i32)]),
        "gpl_server_release_model": (C.c_int, [_vp, i64]),
        "gpl_server_stats": (C.c_int, [_vp] + [C.POINTER(i64)] * 6),
        "gpl_server_destroy": (None, [_vp]),
        "gpl_pred_ctx": (C.c_int, [_vp, _vp, _dp, i64, i32, _dp, _dp]),
        "gpl_draw": (C.c_int, [_vp, _dp, i64, C.c_double, _dp, i64, _dp]),
        "gpl_draw_prior": (C.c_int, [_vp, C.POINTER(CfTerm), i32, _dp, i64, i64, C.c_double, _dp, i64, _dp]),
        "gpl_fitc_fit": (C.c_int, [_vp, C.POINTER(CfTerm), i32, _dp, i64, i64, _dp, i64, _dp, _dp, _dp,
                                   C.POINTER(Lik), C.POINTER(Opts), C.POINTER(_vp), C.POINTER(C.c_double)]),
        "gpl_fitc_pred": (C.c_int, [_vp, _dp, i64, i32, _dp, _dp]),
        "gpl_fitc_draw": (C.c_int, [_vp, _dp, i64, C.c_double, _dp, i64, _dp]),
        "gpl_fitc_vector": (C.c_int, [_vp, i32, _dp]),
        "gpl_fitc_train_var": (C.c_int, [_vp, _dp]),
        "gpl_fitc_info": (C.c_int, [_vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i32), C.POINTER(C.c_double)]),
        "gpl_fitc_free": (None, [_vp]),
        "gpl_draw_terms": (C.c_int, [_vp, C.POINTER(i32), i32, _dp, i64, C.c_double, _dp, i64, _dp]),
        "gpl_pred_terms": (C.c_int, [_vp, C.POINTER(i32), i32, _dp, i64, i32, _dp, _dp]),
        "gpl_model_restore": (C.c_int, [_vp, C.POINTER(CfTerm), i32, _dp, i64, i64, C.POINTER(Lik), C.POINTER(Opts),
                                        _dp, C.POINTER(_dp), C.c_double, i32, C.POINTER(_vp)]),
        "gpl_model_bytes": (i64, [_vp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def lib():
    """Load libgplite_amd.so (raises if it is not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gplite_amd.build`")
            _lib = _sig(C.CDLL(LIB_PATH))
    return _lib


def ctx():
    """The process-wide device context on GPU 0 (raises without a GPU)."""
    global _ctx
    L = lib()
    with _lock:
        if _ctx is None:
            p = _vp()
            rc = L.gpl_ctx_create(0, C.byref(p))
            if rc != GPL_OK:
                raise GplError(rc, L.gpl_last_error().decode())
            _ctx = p
    return _ctx


def check(rc, allow=()):
    if rc != GPL_OK and rc not in allow:
        L = lib()
        raise GplError(rc, L.gpl_last_error().decode(), L.gpl_last_error_minor())
    return rc


def dptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def f64c(a):
    """Column-major float64 copy (R's memory layout)."""
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def profile(enable=None, reset=False):
    L = lib(); c = ctx()
    if enable is not None:
        check(L.gpl_profile_enable(c, int(bool(enable))))
    if reset:
        check(L.gpl_profile_reset(c))
    cap = 128
    names = C.create_string_buffer(32 * cap)
    ms = (C.c_double * cap)(); cnt = (C.c_int64 * cap)(); fl = (C.c_double * cap)(); by = (C.c_double * cap)()
    n = L.gpl_profile_read(c, cap, names, ms, cnt, fl, by)
    out = {}
    for i in range(min(n, cap)):
        nm = names.raw[32 * i:32 * i + 32].split(b"\0", 1)[0].decode()
        out[nm] = {"ms": ms[i], "launches": cnt[i], "flops": fl[i], "bytes": by[i]}
    return out


def device_name():
    buf = C.create_string_buffer(256)
    check(lib().gpl_device_name(ctx(), buf, 256))
    return buf.value.decode()
