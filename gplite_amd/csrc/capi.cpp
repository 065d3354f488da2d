// C-ABI implementation: device context, model handles and the orchestration of
// the dense GP core (fit / energy+gradient / prediction) on one gfx950 GPU.
//
// Reference call chains restated here:
//   gp_fit -> fit_laplace.method_full  R/gp_fit.R:54-60,91-99
//          -> laplace.method_full      R/laplace.R:274-297 (Newton loop)
//          -> laplace_iter.method_full R/laplace.R:13-39,119 (B-form Cholesky chain)
//   gp_fit -> fit_ep.method_full       R/gp_fit.R:150-161 -> ep.method_full R/ep.R:286-330
//   gp_pred -> gp_pred_post.method_full R/gp_pred.R:202-232
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "host_common.h"
#include "model.h"
#include "serve_device.h"

static_assert(gpl::kServeFeatCols == gpl::MAX_FEAT, "the serving footprint counts the device feature columns");

gpl::TraceRange::TraceRange(const char* name) { roctxRangePushA(name); }
gpl::TraceRange::~TraceRange() { roctxRangePop(); }

using namespace gpl;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
static thread_local int64_t g_minor = 0;

namespace gpl {
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
void set_minor(int64_t minor) { g_minor = minor; }
}  // namespace gpl

extern "C" const char* gpl_last_error(void) { return g_err.c_str(); }
extern "C" int64_t gpl_last_error_minor(void) { return g_minor; }

extern "C" int gpl_ctx_create(int32_t device, gpl_ctx** out) {
    if (!out) return fail(GPL_ERR_ARG, "out is NULL");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(GPL_ERR_ARG, "no such device");
    HIP_TRY(hipSetDevice(device));
    {
        const int r = noalloc_init(device);
        if (r) return r;
    }
    gpl_ctx* c = new gpl_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_h, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_t, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_h2, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_a, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_p, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_d, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc((void**)&c->psync, kPanelSyncWords * sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(c->psync, 0, kPanelSyncWords * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc((void**)&c->tsync, kSolveSyncWords * sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(c->tsync, 0, kSolveSyncWords * sizeof(unsigned));
    if (const char* pw = getenv("GPL_PANEL_WAIT_POLLS")) c->pwait = (unsigned)strtoul(pw, nullptr, 10);
    if (e != hipSuccess) {
        delete c;
        return fail(GPL_ERR_HIP, hipGetErrorString(e));
    }
    *out = c;
    return GPL_OK;
}

extern "C" void gpl_ctx_destroy(gpl_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    c->resolve();
    for (auto e : c->pool) hipEventDestroy(e);
    for (auto p : c->scratch)
        if (p) hipFree(p);
    hipStreamSynchronize(c->side);
    hipStreamSynchronize(c->aux);
    hipEventDestroy(c->ev_a);
    hipEventDestroy(c->ev_p);
    hipEventDestroy(c->ev_d);
    hipEventDestroy(c->ev_h);
    hipEventDestroy(c->ev_t);
    hipEventDestroy(c->ev_h2);
    hipStreamDestroy(c->aux);
    if (c->psync) hipFree(c->psync);
    if (c->tsync) hipFree(c->tsync);
    hipStreamDestroy(c->side);
    hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int gpl_device_name(gpl_ctx* c, char* buf, int32_t len) {
    if (!c || !buf || len <= 0) return fail(GPL_ERR_ARG, "bad args");
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, c->device));
    snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return GPL_OK;
}

extern "C" int gpl_profile_enable(gpl_ctx* c, int32_t on) {
    if (!c) return fail(GPL_ERR_ARG, "ctx NULL");
    c->prof = on != 0;
    return GPL_OK;
}

extern "C" int gpl_profile_reset(gpl_ctx* c) {
    if (!c) return fail(GPL_ERR_ARG, "ctx NULL");
    hipStreamSynchronize(c->stream);
    c->resolve();
    for (auto& k : c->classes) k.ms = k.flops = k.bytes = 0, k.launches = 0;
    return GPL_OK;
}

extern "C" int gpl_profile_read(gpl_ctx* c, int32_t cap, char* names, double* total_ms, int64_t* launches,
                                double* flops, double* bytes) {
    if (!c) return fail(GPL_ERR_ARG, "ctx NULL");
    hipStreamSynchronize(c->stream);
    c->resolve();
    int n = (int)c->classes.size();
    for (int i = 0; i < n && i < cap; ++i) {
        const auto& k = c->classes[i];
        if (names) {
            strncpy(names + 32 * i, k.name.c_str(), 31);
            names[32 * i + 31] = 0;
        }
        if (total_ms) total_ms[i] = k.ms;
        if (launches) launches[i] = k.launches;
        if (flops) flops[i] = k.flops;
        if (bytes) bytes[i] = k.bytes;
    }
    return n;
}

extern "C" void gpl_opts_default(gpl_opts* o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->approx = GPL_APPROX_LAPLACE;
    o->maxiter = 30;
    o->tol = 1e-4;
    o->damping = 0.9;
    o->damping_min = 0.1;
    o->quad_order = 11;
    o->precision = GPL_PREC_FP64;
    o->jitter = 1e-6;
    o->refine_iters = 8;
}

// ---------------------------------------------------------------------------
// covariance program preparation
// ---------------------------------------------------------------------------
static int term_ngrad(const gpl_cf_term& t) {
    switch (t.kind) {
        case GPL_CF_CONST:
        case GPL_CF_LIN: return 1;
        case GPL_CF_NN: return 3;
        case GPL_CF_PERIODIC: return 2 + t.n_lscale;
        default: return 1 + t.n_lscale;
    }
}

extern "C" int32_t gpl_grad_len(const gpl_cf_term* terms, int32_t nterms, int32_t lik_kind) {
    int32_t g = 0;
    for (int k = 0; k < nterms; ++k) g += term_ngrad(terms[k]);
    if (lik_kind == GPL_LIK_GAUSSIAN) g += 1;
    return g;
}

int gpl::prepare_prog(const gpl_cf_term* terms, int32_t nterms, int64_t ncols, DevProg& P) {
    memset(&P, 0, sizeof(P));
    if (!terms || nterms < 1 || nterms > GPL_MAX_TERMS) return fail(GPL_ERR_ARG, "nterms must be in [1, 8]");
    int f = 0, g = 0;
    // periodic derivative features are placed after all value features
    int nval = 0;
    for (int k = 0; k < nterms; ++k) {
        const gpl_cf_term& t = terms[k];
        if (t.kind == GPL_CF_CONST) continue;
        nval += (t.kind == GPL_CF_PERIODIC) ? 2 * t.ncols : t.ncols;
    }
    int fd = nval;
    for (int k = 0; k < nterms; ++k) {
        const gpl_cf_term& t = terms[k];
        DevTerm& d = P.t[k];
        if (t.kind < GPL_CF_CONST || t.kind > GPL_CF_PERIODIC) return fail(GPL_ERR_ARG, "unknown cf kind");
        d.kind = t.kind;
        d.op = (k == 0) ? GPL_OP_SUM : t.op;
        d.m2 = t.magn * t.magn;
        if (!(t.magn > 0) && t.kind != GPL_CF_CONST && t.kind != GPL_CF_LIN)
            return fail(GPL_ERR_ARG, "magn must be positive");
        if (t.kind == GPL_CF_CONST) {
            d.nc = d.nf = 0;
        } else {
            if (t.ncols < 1 || t.ncols > GPL_MAX_COLS) return fail(GPL_ERR_ARG, "term ncols must be in [1, 16]");
            if (t.col0 < 0 || t.col0 + t.ncols > ncols) return fail(GPL_ERR_ARG, "term columns out of range");
            d.c0 = t.col0;
            d.nc = t.ncols;
            d.f0 = f;
            d.nf = (t.kind == GPL_CF_PERIODIC) ? 2 * t.ncols : t.ncols;
            f += d.nf;
        }
        if (t.kind == GPL_CF_SEXP || t.kind == GPL_CF_MATERN32 || t.kind == GPL_CF_MATERN52 ||
            t.kind == GPL_CF_PERIODIC) {
            if (t.kind == GPL_CF_PERIODIC) {
                if (t.base_kind != GPL_CF_SEXP && t.base_kind != GPL_CF_MATERN32 && t.base_kind != GPL_CF_MATERN52)
                    return fail(GPL_ERR_ARG, "periodic base must be sexp/matern32/matern52");
                if (!(t.period > 0)) return fail(GPL_ERR_ARG, "period must be positive");
                d.base_kind = t.base_kind;
                d.w = 2.0 * M_PI / t.period;
                d.df0 = fd;
                fd += d.nf;
            }
            if (t.n_lscale != 1 && t.n_lscale != d.nf)
                return fail(GPL_ERR_ARG, "n_lscale must be 1 or the number of (feature) columns");
            d.nl = t.n_lscale;
            for (int c = 0; c < d.nf; ++c) {
                const double l = t.lscale[t.n_lscale == 1 ? 0 : c];
                if (!(l > 0)) return fail(GPL_ERR_ARG, "lscale must be positive");
                d.il2[c] = 1.0 / (l * l);
            }
        }
        if (t.kind == GPL_CF_NN) {
            d.s2 = t.sigma * t.sigma;
            d.b0 = t.sigma0 * t.sigma0 * d.s2;
        }
        d.g0 = g;
        d.ng = term_ngrad(t);
        g += d.ng;
    }
    if (fd > MAX_FEAT) return fail(GPL_ERR_ARG, "too many feature columns (max 64)");
    if (g > GRAD_MAX) return fail(GPL_ERR_ARG, "too many hyperparameters");
    P.nterms = nterms;
    P.nfeat = fd;
    P.ngrad = g;
    return GPL_OK;
}

// Outer row-panel height of the forward-substitution inverse (trtri_blocked
// without scratch): 256 rows in 64-row steps.
static const int kOuterNB = 256;
// Outer panel width of the two-level right-looking Cholesky drivers: each panel
// is factored in one launch (k_potrf_panel: 64-column blocks inside), then the
// trailing matrix receives one rank-kPanelNB update.  512 columns measured 4 %
// faster than 256 at n = 8192 (FP64 and mixed) and 3 % at n = 4096.
static const int kPanelNB = 512;
static_assert(kPanelNB / 64 <= kPanelMaxBlocks, "panel wider than the panel kernel's flags");

// X = L^-1 by recursive doubling from the 64x64 diagonal-block inverses Dinv.
// At level b (b = 64, 128, ...) the b-blocks pair up: [[X11, 0], [L21, X22]]
// with X11 and X22 already inverted has the inverse [[X11, 0], [X21, X22]],
// X21 = -X22 (L21 X11).  The pairs of one level are independent, so each level
// is two batched GEMMs (plus one more pair when the last is shorter): with
// N/64 diagonal blocks the dependent chain is 2 log2(N/64) launches, against
// two launches per 64 rows for forward substitution.  Both triangular operands
// skip their zero tiles (tile modes 4 and 3).  T (N x N, ld N) holds the
// products L21 X11 at the positions of X21.  FP64 (T = double) or FP32 storage
// (float: FP32 MFMA with sums folded into FP64 every 64 k, FP64 leaves rounded).
//
// The work splits at the top level bt = doubling_top(N): the inverse of the
// diagonal block [0, bt) and T = L21 X11 below it (dbl_front) need only the
// factor's first bt columns, so the blocked Cholesky can run them on the
// context's third stream beside its remaining panels; the inverse of [bt, N)
// and X21 = -X22 T (dbl_back) follow the factorization.
static void dbl_leaves(hipStream_t s, double* X, int64_t N, const double* Dinv, int64_t c0, int64_t n) {
    trtri_leaves_f64(s, X, N, Dinv, c0, n);
}
static void dbl_leaves(hipStream_t s, float* X, int64_t N, const double* Dinv, int64_t c0, int64_t n) {
    trtri_leaves_f32_batched(s, X, N, Dinv, 1, 0, 0, c0, n);
}
static void dbl_gemm(hipStream_t s, int64_t M, int64_t Nn, int64_t K, double alpha, const double* A, const double* B,
                     double* C, int64_t ld, int tiles, int64_t batch, int64_t step, double beta = 0.0) {
    if (batch == 1) gemm2_f64(s, 0, 0, M, Nn, K, alpha, A, ld, B, ld, beta, C, ld, tiles, 0);
    else gemm2_f64_batched(s, 0, 0, M, Nn, K, alpha, A, ld, B, ld, beta, C, ld, tiles, (int)batch, step, step, step);
}
static void dbl_gemm(hipStream_t s, int64_t M, int64_t Nn, int64_t K, double alpha, const float* A, const float* B,
                     float* C, int64_t ld, int tiles, int64_t batch, int64_t step, double beta = 0.0) {
    gemm_f32_batched(s, 0, 0, M, Nn, K, (float)alpha, A, ld, B, ld, (float)beta, C, ld, tiles, 0, (int)batch, step,
                     step, step, kFoldInverse);
}

// the leaves (columns r0 .. r0 + n) and the levels inside the diagonal block
// [r0, r0 + n); r0 is 0 or bt, a multiple of 2b for every level run
template <typename T>
static void dbl_levels(gpl_ctx* c, hipStream_t s, const T* L, const double* Dinv, int64_t N, T* X, T* Tm, int64_t r0,
                       int64_t n) {
    const bool f32 = sizeof(T) == 4;
    {
        Prof p(c, f32 ? "trtri_leaves_f32" : "trtri_leaves", 0, sizeof(T) * (double)N * n / 2, s);
        dbl_leaves(s, X, N, Dinv, r0, n);
    }
    const int64_t o0 = r0 * (N + 1);
    L += o0;
    X += o0;
    Tm += o0;
    for (int64_t b = 64; b < n; b *= 2) {
        const int64_t full = n / (2 * b);  // pairs whose right block has b rows
        const int64_t r = std::min<int64_t>(b, n - full * 2 * b - b);  // rows of a last, shorter right block
        const int64_t step = 2 * b * (N + 1);
        Prof p(c, f32 ? "trtri_level_f32" : "trtri_level",
               2.0 * (double)b * b * b * full + (r > 0 ? (double)r * b * (b + r) : 0.0), 0, s);
        if (full > 0) {
            dbl_gemm(s, b, b, b, 1.0, L + b, X, Tm + b, N, 4, full, step);
            dbl_gemm(s, b, b, b, -1.0, X + b * (N + 1), Tm + b, X + b, N, 3, full, step);
        }
        if (r > 0) {
            const int64_t o = full * step;  // the last pair's top-left corner
            dbl_gemm(s, r, b, b, 1.0, L + o + b, X + o, Tm + o + b, N, 4, 1, 0);
            dbl_gemm(s, r, b, r, -1.0, X + o + b * (N + 1), Tm + o + b, X + o + b, N, 3, 1, 0);
        }
    }
}

template <typename T>
static void dbl_front(gpl_ctx* c, hipStream_t s, const T* L, const double* Dinv, int64_t N, T* X, T* Tm, int64_t bt) {
    dbl_levels(c, s, L, Dinv, N, X, Tm, 0, bt);
    Prof p(c, sizeof(T) == 4 ? "trtri_level_f32" : "trtri_level", (double)(N - bt) * bt * bt, 0, s);
    dbl_gemm(s, N - bt, bt, bt, 1.0, L + bt, X, Tm + bt, N, 4, 1, 0);
}

template <typename T>
static void dbl_back(gpl_ctx* c, hipStream_t s, const T* L, const double* Dinv, int64_t N, T* X, T* Tm, int64_t bt) {
    dbl_levels(c, s, L, Dinv, N, X, Tm, bt, N - bt);
    Prof p(c, sizeof(T) == 4 ? "trtri_level_f32" : "trtri_level", (double)(N - bt) * (N - bt) * bt, 0, s);
    dbl_gemm(s, N - bt, bt, N - bt, -1.0, X + bt * (N + 1), Tm + bt, X + bt, N, 3, 1, 0);
}

template <typename T>
static void trtri_doubling(gpl_ctx* c, hipStream_t s, const T* L, const double* Dinv, int64_t N, T* X, T* Tm) {
    const int64_t bt = doubling_top(N);
    if (bt == 0) {
        dbl_levels(c, s, L, Dinv, N, X, Tm, 0, N);  // one 64-block: the leaf alone
        return;
    }
    dbl_front(c, s, L, Dinv, N, X, Tm, bt);
    dbl_back(c, s, L, Dinv, N, X, Tm, bt);
}

// The blocked Cholesky's fork and join of L^-1 (X, Tm given).  The front half
// (the inverse of the leading bt x bt block and T = L21 X11) runs on the
// context's third stream once the panel holding column bt - 1 is factored on
// the main stream.  The back half X22 = inv(L22) splits the same way at bt2:
// once column bt2 - 1 is factored, the third stream also forms its leading
// block A = inv(L(bt:bt2, bt:bt2)), T2 = L(bt2:N, bt:bt2) A and the top rows of
// X21 = -X22 T, which need only A (X21(bt:bt2, :) = -A T(bt:bt2, :)).  What
// follows the factor is then C = inv(L(bt2:N, bt2:N)), B = -C T2 and the bottom
// rows X21(bt2:N, :) = -B T(bt:bt2, :) - C T(bt2:N, :).
template <typename T>
struct InverseFork {
    gpl_ctx* c;
    const T* L;
    const double* Dinv;
    int64_t N, bt, bt2;
    T *X, *Tm;
    bool forked = false, forked2 = false;
    InverseFork(gpl_ctx* c_, const T* L_, const double* Dinv_, int64_t N_, T* X_, T* Tm_)
        : c(c_), L(L_), Dinv(Dinv_), N(N_), bt(X_ ? doubling_top(N_) : 0), X(X_), Tm(Tm_) {
        bt2 = (bt > 0 && N - bt > 64) ? bt + doubling_top(N - bt) : 0;
    }
    // after the panel ending at column `done` was queued on the main stream
    void after_panel(int64_t done) {
        if (bt == 0) return;
        if (!forked && done >= bt) {
            hipEventRecord(c->ev_h, c->stream);
            hipStreamWaitEvent(c->aux, c->ev_h, 0);
            dbl_front(c, c->aux, L, Dinv, N, X, Tm, bt);
            hipEventRecord(c->ev_t, c->aux);
            forked = true;
        }
        if (forked && !forked2 && bt2 > 0 && done >= bt2) {
            hipEventRecord(c->ev_h2, c->stream);
            hipStreamWaitEvent(c->aux, c->ev_h2, 0);
            const int64_t h2 = bt2 - bt, r2 = N - bt2;
            hipStream_t a = c->aux;
            dbl_levels(c, a, L, Dinv, N, X, Tm, bt, h2);
            Prof p(c, sizeof(T) == 4 ? "trtri_level_f32" : "trtri_level", (double)r2 * h2 * h2 + (double)h2 * h2 * bt, 0, a);
            dbl_gemm(a, r2, h2, h2, 1.0, L + bt2 + bt * N, X + bt * (N + 1), Tm + bt2 + bt * N, N, 4, 1, 0);
            dbl_gemm(a, h2, bt, h2, -1.0, X + bt * (N + 1), Tm + bt, X + bt, N, 3, 1, 0);
            hipEventRecord(c->ev_t, a);
            forked2 = true;
        }
    }
    void finish() {
        if (!X) return;
        if (bt == 0) {
            dbl_levels(c, c->stream, L, Dinv, N, X, Tm, 0, N);
            return;
        }
        hipStreamWaitEvent(c->stream, c->ev_t, 0);
        if (!forked2) {
            dbl_back(c, c->stream, L, Dinv, N, X, Tm, bt);
            return;
        }
        hipStream_t s = c->stream;
        const int64_t h2 = bt2 - bt, r2 = N - bt2;
        dbl_levels(c, s, L, Dinv, N, X, Tm, bt2, r2);
        Prof p(c, sizeof(T) == 4 ? "trtri_level_f32" : "trtri_level",
               (double)r2 * r2 * h2 + 2.0 * r2 * h2 * bt + (double)r2 * r2 * bt, 0, s);
        dbl_gemm(s, r2, h2, r2, -1.0, X + bt2 * (N + 1), Tm + bt2 + bt * N, X + bt2 + bt * N, N, 3, 1, 0);
        dbl_gemm(s, r2, bt, h2, -1.0, X + bt2 + bt * N, Tm + bt, X + bt2, N, 0, 1, 0);
        dbl_gemm(s, r2, bt, r2, -1.0, X + bt2 * (N + 1), Tm + bt2, X + bt2, N, 3, 1, 0, 1.0);
    }
};

// Factor the outer panel of columns [K0, K0+W) (all rows >= K0) on stream st:
// one launch, a workgroup per 64-row block (k_potrf_panel).
static void potrf_panel(gpl_ctx* c, hipStream_t st, double* L, int64_t N, int64_t K0, int64_t W, double* Dinv,
                        double* ldiag, int* info, double* rhs) {
    const int64_t nrb = (N - K0) / 64;
    if (++c->pepoch == 0) c->pepoch = 1;
    {
        Prof p(c, "potrf_panel", (double)W * W * W / 3 + 2.0 * (N - K0 - W) * W * W, 8.0 * (N - K0) * W * 3, st);
        potrf_panel_f64(st, L, N, K0, (int)(W / 64), nrb, Dinv, ldiag, info, c->psync, c->pepoch, c->pbase, c->pwait,
                        rhs);
    }
    c->pbase += (unsigned)nrb;
}

// Two-level blocked right-looking Cholesky with one-panel look-ahead: after the
// rank-NB update of the next panel's columns, the main stream factors that
// panel (one launch) while the side stream applies the update to the rest of
// the trailing matrix.  The panel launch is queued first: its workgroups (one
// per 64-row block, each needing half a CU's registers) are placed before the
// update's GEMM fills the CUs, instead of waiting for it to drain.  With X and
// Tm (N x N scratch) it also forms X = L^-1 (recursive doubling, its front half
// beside the remaining panels); with rhs it overwrites that vector with L^-1 rhs.
void gpl::potrf_blocked(gpl_ctx* c, double* L, int64_t N, double* Dinv, double* ldiag, int* info, double* X,
                        double* Tm, double* rhs) {
    hipStream_t s = c->stream, s2 = c->side;
    InverseFork<double> inv(c, L, Dinv, N, X, Tm);
    potrf_panel(c, s, L, N, 0, std::min<int64_t>(kPanelNB, N), Dinv, ldiag, info, rhs);
    inv.after_panel(std::min<int64_t>(kPanelNB, N));
    for (int64_t K0 = 0; K0 < N; K0 += kPanelNB) {
        const int64_t W = std::min<int64_t>(kPanelNB, N - K0);
        const int64_t K1 = K0 + W, R = N - K1;
        if (R <= 0) break;
        const int64_t W1 = std::min<int64_t>(kPanelNB, R);
        const double* P = L + K1 + K0 * N;  // rows >= K1 of the factored panel
        {
            Prof p(c, "syrk_update", 2.0 * R * W1 * W, 8.0 * (2.0 * R * W1 + R * W), s);
            gemm2_f64(s, 0, 1, R, W1, W, -1.0, P, N, P, N, 1.0, L + K1 * (N + 1), N, 0, 0);
        }
        hipEventRecord(c->ev_a, s);
        potrf_panel(c, s, L, N, K1, W1, Dinv, ldiag, info, rhs);
        inv.after_panel(K1 + W1);
        const int64_t R2 = R - W1;
        if (R2 > 0) {
            hipStreamWaitEvent(s2, c->ev_a, 0);
            const double* P2 = L + (K1 + W1) + K0 * N;
            {
                Prof p(c, "syrk_update", (double)R2 * R2 * W, 8.0 * ((double)R2 * R2 + 2.0 * R2 * W), s2);
                gemm2_f64(s2, 0, 1, R2, R2, W, -1.0, P2, N, P2, N, 1.0, L + (K1 + W1) * (N + 1), N, 1, 0);
            }
            hipEventRecord(c->ev_p, s2);
            hipStreamWaitEvent(s, c->ev_p, 0);
        }
    }
    inv.finish();
}

// Rows [I0, I0+H) of X = L^-1: 64-row steps inside the panel, on stream st.
static void trtri_panel(gpl_ctx* c, hipStream_t st, const double* L, const double* Dinv, int64_t N, int64_t I0,
                        int64_t H, double* X) {
    for (int64_t r0 = I0; r0 < I0 + H; r0 += 64) {
        const int64_t i = r0 / 64, Wc = r0 + 64;  // columns 0 .. r0+63 are non-zero
        {
            Prof p(c, "trtri_diag", 2.0 * 64 * 64 * Wc, 16.0 * 64 * Wc, st);
            gemm2_f64(st, 0, 0, 64, Wc, 64, 1.0, Dinv + i * 64 * 64, 64, X + r0, N, 0.0, X + r0, N, 0, 0);
        }
        const int64_t Mr = I0 + H - (r0 + 64);
        if (Mr > 0) {
            Prof p(c, "trtri_panel", 2.0 * Mr * Wc * 64, 16.0 * Mr * Wc, st);
            gemm2_f64(st, 0, 0, Mr, Wc, 64, -1.0, L + (r0 + 64) + r0 * N, N, X + r0, N, 1.0, X + r0 + 64, N, 0, 0);
        }
    }
}

// X = L^-1 by blocked forward substitution against the identity, two levels with
// look-ahead: the next row panel is updated first and solved on the side stream
// while the main stream updates the remaining rows.  With scratch T (N x N) the
// recursive doubling above runs instead.
void gpl::trtri_blocked(gpl_ctx* c, const double* L, const double* Dinv, int64_t N, double* X, double* T) {
    if (T) {
        trtri_doubling(c, c->stream, L, Dinv, N, X, T);
        return;
    }
    hipStream_t s = c->stream, s2 = c->side;
    set_identity_f64(s, X, N, N);
    trtri_panel(c, s, L, Dinv, N, 0, std::min<int64_t>(kOuterNB, N), X);
    for (int64_t I0 = 0; I0 < N; I0 += kOuterNB) {
        const int64_t H = std::min<int64_t>(kOuterNB, N - I0);
        const int64_t I1 = I0 + H, R = N - I1;
        if (R <= 0) break;
        const int64_t H1 = std::min<int64_t>(kOuterNB, R);
        const int64_t Wc = I0 + H;
        {
            Prof p(c, "trtri_update", 2.0 * H1 * Wc * H, 8.0 * (2.0 * H1 * Wc + Wc * H), s);
            gemm2_f64(s, 0, 0, H1, Wc, H, -1.0, L + I1 + I0 * N, N, X + I0, N, 1.0, X + I1, N, 0, 0);
        }
        hipEventRecord(c->ev_a, s);
        hipStreamWaitEvent(s2, c->ev_a, 0);
        trtri_panel(c, s2, L, Dinv, N, I1, H1, X);
        hipEventRecord(c->ev_p, s2);
        const int64_t R2 = R - H1;
        if (R2 > 0) {
            Prof p(c, "trtri_update", 2.0 * R2 * Wc * H, 8.0 * (2.0 * R2 * Wc + Wc * H), s);
            gemm2_f64(s, 0, 0, R2, Wc, H, -1.0, L + (I1 + H1) + I0 * N, N, X + I0, N, 1.0, X + I1 + H1, N, 0, 0);
        }
        hipStreamWaitEvent(s, c->ev_p, 0);
    }
}

// ---------------------------------------------------------------------------
// Mixed precision (GPL_PREC_MIXED): FP32 storage and FP32 MFMA updates; every
// 64x64 diagonal block is factored and inverted in FP64 and the panel solve
// against it runs in FP64 arithmetic; the log-determinant accumulates the FP64
// diagonal-block pivots.
// ---------------------------------------------------------------------------
// Factor the outer panel of columns [K0, K0+W) (all rows >= K0), FP32 storage
// with the FP64 diagonal tiles Dacc: one launch (k_potrf_panel<float>).
static void potrf32_panel(gpl_ctx* c, hipStream_t st, float* L, int64_t N, int64_t K0, int64_t W, double* Dinv,
                          double* ldiag, int* info, double* Dacc) {
    const int64_t nrb = (N - K0) / 64;
    if (++c->pepoch == 0) c->pepoch = 1;
    {
        Prof p(c, "potrf_panel_mixed", (double)W * W * W / 3 + 2.0 * (N - K0 - W) * W * W, 4.0 * (N - K0) * W * 3, st);
        potrf_panel_mixed(st, L, N, K0, (int)(W / 64), nrb, Dinv, ldiag, info, c->psync, c->pepoch, c->pbase, Dacc,
                          c->pwait);
    }
    c->pbase += (unsigned)nrb;
}

// Dacc: FP64 copies of the diagonal tiles (N x 64), filled by the build.  The
// look-ahead as in potrf_blocked: the next panel on the main stream, queued
// ahead of the rest of the trailing update on the side stream.
static void potrf32_blocked(gpl_ctx* c, float* L, int64_t N, double* Dinv, double* ldiag, int* info, double* Dacc,
                            float* X = nullptr, float* Tm = nullptr) {
    hipStream_t s = c->stream, s2 = c->side;
    InverseFork<float> inv(c, L, Dinv, N, X, Tm);
    potrf32_panel(c, s, L, N, 0, std::min<int64_t>(kPanelNB, N), Dinv, ldiag, info, Dacc);
    inv.after_panel(std::min<int64_t>(kPanelNB, N));
    for (int64_t K0 = 0; K0 < N; K0 += kPanelNB) {
        const int64_t W = std::min<int64_t>(kPanelNB, N - K0);
        const int64_t K1 = K0 + W, R = N - K1;
        if (R <= 0) break;
        const int64_t W1 = std::min<int64_t>(kPanelNB, R);
        const float* P = L + K1 + K0 * N;
        // the FP64 diagonal tiles of the next panel on the side stream, beside
        // its FP32 update (a few workgroups, latency-bound)
        hipEventRecord(c->ev_d, s);
        hipStreamWaitEvent(s2, c->ev_d, 0);
        {
            Prof p(c, "dacc_f64", 2.0 * 1 * 4096.0 * W * (W1 / 64), 0, s2);
            dacc_update(s2, Dacc + (K1 / 64) * 64 * 64, P, N, W1 / 64, W, 1, 0, 0);
        }
        hipEventRecord(c->ev_d, s2);
        {
            Prof p(c, "syrk_update_f32", 2.0 * R * W1 * W, 4.0 * (2.0 * R * W1 + R * W), s);
            gemm_f32(s, 0, 1, R, W1, W, -1.f, P, N, P, N, 1.f, L + K1 * (N + 1), N, 0, 0);
        }
        hipStreamWaitEvent(s, c->ev_d, 0);
        hipEventRecord(c->ev_a, s);
        potrf32_panel(c, s, L, N, K1, W1, Dinv, ldiag, info, Dacc);
        inv.after_panel(K1 + W1);
        const int64_t R2 = R - W1;
        if (R2 > 0) {
            hipStreamWaitEvent(s2, c->ev_a, 0);
            const float* P2 = L + (K1 + W1) + K0 * N;
            {
                Prof p(c, "syrk_update_f32", (double)R2 * R2 * W, 4.0 * ((double)R2 * R2 + 2.0 * R2 * W), s2);
                gemm_f32(s2, 0, 1, R2, R2, W, -1.f, P2, N, P2, N, 1.f, L + (K1 + W1) * (N + 1), N, 1, 0);
            }
            {
                Prof p(c, "dacc_f64", 2.0 * 1 * 4096.0 * W * (R2 / 64), 0, s2);
                dacc_update(s2, Dacc + ((K1 + W1) / 64) * 64 * 64, P2, N, R2 / 64, W, 1, 0, 0);
            }
            hipEventRecord(c->ev_p, s2);
            hipStreamWaitEvent(s, c->ev_p, 0);
        }
    }
    inv.finish();
}

// alpha = sv .* L^-T L^-1 (sv .* z)   (R/laplace.R:29 with C = diag(sqrt V) L)
// alpha = S B^-1 S z = S L^-T L^-1 S z with the factorization carrying the
// forward solve: W1 = S z before potrf_blocked(..., rhs = W1), which leaves
void gpl::trsv_lower(gpl_ctx* c, hipStream_t s, const double* L, int64_t ld, const double* Dinv, int64_t N, double* b,
                     double* x, int trans, int* info) {
    const int64_t T = N / 64;
    if (T > kSolveMaxBlocks) {
        trsv_lower_f64(s, L, ld, Dinv, N, b, x, trans);
        return;
    }
    if (++c->pepoch == 0) c->pepoch = 1;
    trsv_persist_f64(s, L, ld, Dinv, N, b, x, trans, c->tsync, c->pepoch, c->tbase, info, c->pwait);
    c->tbase += (unsigned)T;
}

// L^-1 S z in W1; then alpha = S L^-T W1.
static double* alpha_rhs(gpl_model* m, const double* z) {
    hipStream_t s = m->ctx->stream;
    double* w1 = m->d(B_W1);
    fill_f64(s, w1, m->N, 0.0);
    vec_mul(s, m->n, m->d(B_SV), z, w1);
    return w1;
}

static void solve_alpha_back(gpl_ctx* c, gpl_model* m, double* alpha) {
    hipStream_t s = c->stream;
    const int64_t N = m->N;
    double* w1 = m->d(B_W1);
    double* w2 = m->d(B_W2);
    {
        Prof p(c, "trsv", 1.0 * N * N, 8.0 * N * N / 2);
        trsv_lower(c, s, m->d(B_L), N, m->d(B_DINV), N, w1, w2, 1, m->b[B_INFO].as<int>());
    }
    fill_f64(s, alpha, N, 0.0);
    vec_mul(s, m->n, m->d(B_SV), w2, alpha);
}

int gpl::wait_timeout_fail() {
    return fail(GPL_ERR_HIP, "blocked Cholesky or triangular solve: a wait between the workgroups of a launch timed out");
}

// A factorisation's info word: 0, the first failing row + 1, or kPanelWaitTimeout.
static int pivot_fail(int info) {
    if (info == kPanelWaitTimeout) return gpl::wait_timeout_fail();
    g_minor = info;
    return fail(GPL_ERR_NOT_PD, "the leading minor of order " + std::to_string(info) + " is not positive");
}

static int check_info(gpl_model* m) {
    int info = 0;
    HIP_TRY(hipMemcpyAsync(&info, m->b[B_INFO].p, sizeof(int), hipMemcpyDeviceToHost, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    if (info > 0) return pivot_fail(info);
    return GPL_OK;
}

static int alloc_model(gpl_model* m, int64_t n, int64_t ncols, bool need_k, bool mixed);

int gpl::model_alloc(gpl_model* m, int64_t n, int64_t ncols, bool mixed) { return alloc_model(m, n, ncols, false, mixed); }

static int alloc_model(gpl_model* m, int64_t n, int64_t ncols, bool need_k, bool mixed) {
    const int64_t N = pad64(n);
    m->n = n;
    m->N = N;
    m->ncols = ncols;
    const size_t vec = (size_t)N * 8;
    // the mixed form keeps its factor in FP32
    const size_t mat = (size_t)N * N * (mixed ? 4 : 8);
    ENSURE(m, B_X, (size_t)n * ncols * 8);
    ENSURE(m, B_F, (size_t)N * MAX_FEAT * 8);
    ENSURE(m, B_Y, vec);
    ENSURE(m, B_TR, vec);
    ENSURE(m, B_OFF, vec);
    ENSURE(m, B_L, mat);
    ENSURE(m, B_DINV, (size_t)N * 64 * 8);
    ENSURE(m, B_LDIAG, vec);
    if (need_k) ENSURE(m, B_K, (size_t)N * N * 8);  // K stays FP64 in both forms
    ENSURE(m, B_PROG, sizeof(DevProg));
    for (int i : {B_Z, B_V, B_SV, B_ALPHA, B_FMEAN, B_FVAR, B_CAVM, B_CAVV, B_LOGZS, B_TERM, B_FOLD, B_FVOLD, B_ZOLD,
                  B_POLD, B_W1, B_W2, B_W3, B_W4})
        ENSURE(m, i, vec);
    ENSURE(m, B_SCAL, 256 * 8);
    ENSURE(m, B_KMV, (size_t)N * std::max(kmatvec_chunks(N), symv_slots(N, 1)) * 8);
    ENSURE(m, B_GH, 64 * 8);
    ENSURE(m, B_INFO, 64);
    return GPL_OK;
}

static int upload_vec(gpl_model* m, int idx, const double* h, int64_t n) {
    hipStream_t s = m->ctx->stream;
    fill_f64(s, m->d(idx), m->N, 0.0);
    if (h) HIP_TRY(hipMemcpyAsync(m->d(idx), h, n * 8, hipMemcpyHostToDevice, s));
    return GPL_OK;
}

// Read small device scalars (one synchronisation).
static int read_scalars(gpl_model* m, double* out, int k) {
    HIP_TRY(hipMemcpyAsync(out, m->d(B_SCAL), k * 8, hipMemcpyDeviceToHost, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return GPL_OK;
}

// One Gaussian "Newton step" (exactly one: R/laplace.R:281-283).
static int fit_gaussian(gpl_model* m) {
    gpl_ctx* c = m->ctx;
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N;
    const double sigma = m->lik.sigma;
    const DevProg* pd = m->b[B_PROG].as<DevProg>();
    {
        Prof p(c, "build_B", 0, 4.0 * N * (N + 64));
        launch_build_sym_f64(s, pd, m->prog, m->d(B_F), n, N, m->opts.jitter, 1, 1.0 / sigma, nullptr, m->d(B_L), N);
    }
    lik_pseudodata_la(s, GPL_LIK_GAUSSIAN, 0, sigma, n, nullptr, m->d(B_Y), nullptr, nullptr, m->d(B_Z), m->d(B_V),
                      m->d(B_SV), m->b[B_INFO].as<int>() + 1);
    const bool inv = m->opts.flags & GPL_FLAG_INVERSE;
    if (inv) {
        // X = L^-1 with the factor; B^-1's buffer (which the gradient fills next)
        // is the inverse's scratch
        ENSURE(m, B_XINV, (size_t)N * N * 8);
        ENSURE(m, B_BINV, (size_t)N * N * 8);
    }
    potrf_blocked(c, m->d(B_L), N, m->d(B_DINV), m->d(B_LDIAG), m->b[B_INFO].as<int>(), inv ? m->d(B_XINV) : nullptr,
                  inv ? m->d(B_BINV) : nullptr, inv ? nullptr : alpha_rhs(m, m->d(B_Z)));
    if (inv) {
        // alpha = S X^T X S z (S = 1/sigma)
        m->have_xinv = true;
        Prof p(c, "trmv", 4.0 * N * N / 2, 8.0 * N * N);
        fill_f64(s, m->d(B_W1), N, 0.0);
        vec_mul(s, n, m->d(B_SV), m->d(B_Z), m->d(B_W1));
        trmv_lower_f64_chunked(s, m->d(B_XINV), N, N, m->d(B_W1), m->d(B_W2), 0, m->d(B_KMV), kmatvec_chunks(N));
        trmv_lower_f64_chunked(s, m->d(B_XINV), N, N, m->d(B_W2), m->d(B_W1), 1, m->d(B_KMV), kmatvec_chunks(N));
        fill_f64(s, m->d(B_ALPHA), N, 0.0);
        vec_mul(s, n, m->d(B_SV), m->d(B_W1), m->d(B_ALPHA));
    } else {
        solve_alpha_back(c, m, m->d(B_ALPHA));
    }
    {
        Prof p(c, "kmatvec", 0, 0);
        launch_kmatvec(s, pd, m->prog, m->d(B_F), n, N, m->opts.jitter, m->d(B_ALPHA), m->d(B_FMEAN), m->d(B_KMV),
                       kmatvec_chunks(N));
    }
    lik_loglik(s, GPL_LIK_GAUSSIAN, 0, sigma, n, m->d(B_FMEAN), m->d(B_Y), nullptr,
               m->has_offset ? m->d(B_OFF) : nullptr, m->d(B_W3));
    double* sc = m->d(B_SCAL);
    reduce_sum(s, m->d(B_W3), n, sc + 0);
    reduce_dot(s, m->d(B_FMEAN), m->d(B_ALPHA), n, sc + 1);
    reduce_sum(s, m->d(B_LDIAG), n, sc + 2);
    int r = check_info(m);
    if (r) return r;
    double h[3];
    r = read_scalars(m, h, 3);
    if (r) return r;
    // log_lik - 1/2 f^T alpha - 1/2 (logdet C - logdet V), logdet C - logdet V = 2 sum log L_ii
    m->logZ = h[0] - 0.5 * h[1] - h[2];
    m->iters = 1;
    return GPL_OK;
}

// Iterative refinement of the mixed form must contract: cond(B) * 2^-24 well
// below 1.  A last relative correction above kRefineTol, a correction that grows
// between sweeps while still above it, or a non-finite one means the FP32
// inverse factor cannot carry the solve, and the fit is redone in FP64
// (internal status kRefineFail, never returned to the caller).
static const double kRefineTol = 1e-9;
static const int kRefineFail = 1000;
// A sweep whose correction is below kRefineStop (relative to the solution) is
// the last: alpha is then good to about that relative size, and the residual
// against the FP64 B (its own rounding about n * 2^-53) leaves little below it
// for further sweeps to gain.
static const double kRefineStop = 1e-12;

static bool refine_diverging(int sweep, double rel, double prev) {
    if (!std::isfinite(rel)) return true;
    return sweep >= 2 && rel >= prev && rel > kRefineTol;
}

// Gaussian fit in the mixed-precision form.  alpha = S B^-1 S y is obtained by
// iterative refinement: corrections come from the FP32 inverse factor
// (d = X^T X r), residuals r = S y - B u are formed in FP64 from the FP64 copy
// of B written by the build, so the converged alpha and f = K alpha carry FP64
// accuracy.
static int fit_gaussian_mixed(gpl_model* m) {
    gpl_ctx* c = m->ctx;
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N;
    const double sigma = m->lik.sigma;
    const DevProg* pd = m->b[B_PROG].as<DevProg>();
    float* L = m->b[B_L].as<float>();
    ENSURE(m, B_XINV, (size_t)N * N * 4);
    ENSURE(m, B_DACC, (size_t)N * 64 * 8);
    ENSURE(m, B_B64, (size_t)N * N * 8);
    const int nch = kmatvec_chunks(N);
    ENSURE(m, B_KMV, (size_t)N * std::max(nch, symv_slots(N, 1)) * 8);
    float* X = m->b[B_XINV].as<float>();
    double* B64 = m->d(B_B64);
    {
        Prof p(c, "build_B_f32", 0, 12.0 * N * (N + 64) / 2);
        launch_build_sym_f32(s, pd, m->prog, m->d(B_F), n, N, m->opts.jitter, 1, 1.0 / sigma, nullptr, L, N,
                             m->d(B_DACC), B64);
    }
    lik_pseudodata_la(s, GPL_LIK_GAUSSIAN, 0, sigma, n, nullptr, m->d(B_Y), nullptr, nullptr, m->d(B_Z), m->d(B_V),
                      m->d(B_SV), m->b[B_INFO].as<int>() + 1);
    // B^-1's buffer (filled below) is the inverse's scratch first
    ENSURE(m, B_BINV, (size_t)N * N * 4);
    potrf32_blocked(c, L, N, m->d(B_DINV), m->d(B_LDIAG), m->b[B_INFO].as<int>(), m->d(B_DACC), X,
                    m->b[B_BINV].as<float>());
    m->have_xinv = true;
    // B^-1 = X^T X (kept for the gradient) and the first-order log|B| correction
    // for having factored fl32(B) rather than B
    const int64_t T = N / 64, ntiles = T * (T + 1) / 2;
    ENSURE(m, B_PART, (size_t)ntiles * std::max(1, m->prog.ngrad) * 8);
    float* Bi = m->b[B_BINV].as<float>();
    {
        Prof p(c, "lauum_f32", (double)N * N * N / 3, 4.0 * N * N);
        gemm_f32(s, 1, 0, N, N, N, 1.f, X, N, X, N, 0.f, Bi, N, 2, 0, kFoldInverse);
    }
    m->have_binv = true;
    {
        Prof p(c, "round_corr", 0, 12.0 * N * (N + 64) / 2);
        launch_round_corr(s, B64, Bi, N, m->d(B_PART), 1, 0, 0);
    }
    reduce_sum(s, m->d(B_PART), ntiles, m->d(B_SCAL) + 3);
    double* rhs = m->d(B_TERM);
    double* u = m->d(B_W1);
    double* r = m->d(B_W2);
    double* d = m->d(B_W3);
    double* t1 = m->d(B_CAVM);
    double* Ku = m->d(B_FMEAN);
    fill_f64(s, rhs, N, 0.0);
    vec_mul(s, n, m->d(B_SV), m->d(B_Z), rhs);  // S y
    fill_f64(s, u, N, 0.0);
    HIP_TRY(hipMemcpyAsync(r, rhs, N * 8, hipMemcpyDeviceToDevice, s));
    const int maxit = m->opts.refine_iters > 0 ? m->opts.refine_iters : 8;
    int it = 0;
    double rel = INFINITY, prev = INFINITY;
    for (it = 1; it <= maxit; ++it) {
        {
            Prof p(c, "refine_solve_f32", 4.0 * N * N / 2, 4.0 * N * N);
            trmv_lower_f32(s, X, N, N, r, t1, 0, m->d(B_KMV), nch);
            trmv_lower_f32(s, X, N, N, t1, d, 1, m->d(B_KMV), nch);
        }
        vec_axpby(s, n, 1.0, d, 1.0, u, u);
        // r = S y - B u with the FP64 B (t1 = B u)
        {
            Prof p(c, "symv_b64", 2.0 * N * N, 8.0 * N * (N + 64) / 2);
            symv_lt_f64(s, B64, N, u, t1, m->d(B_KMV), nch, 1, 0, 0);
        }
        vec_axpby(s, n, 1.0, rhs, -1.0, t1, r);
        rel_maxabs(s, d, u, n, m->d(B_SCAL) + 8);
        double h;
        HIP_TRY(hipMemcpyAsync(&h, m->d(B_SCAL) + 8, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        rel = h;
        if (rel < kRefineStop) break;
        if (refine_diverging(it, rel, prev)) break;
        prev = rel;
    }
    m->refine_iters_done = std::min(it, maxit);
    m->refine_rel = rel;
    if (!(rel <= kRefineTol)) {
        int rr = check_info(m);
        return rr ? rr : kRefineFail;
    }
    fill_f64(s, m->d(B_ALPHA), N, 0.0);
    vec_mul(s, n, m->d(B_SV), u, m->d(B_ALPHA));
    // f = (K + jitter I) alpha = sigma (B u - u), B u from the last sweep (u updated before it)
    vec_axpby(s, n, sigma, t1, -sigma, u, Ku);
    lik_loglik(s, GPL_LIK_GAUSSIAN, 0, sigma, n, m->d(B_FMEAN), m->d(B_Y), nullptr,
               m->has_offset ? m->d(B_OFF) : nullptr, m->d(B_W3));
    double* sc = m->d(B_SCAL);
    reduce_sum(s, m->d(B_W3), n, sc + 0);
    reduce_dot(s, m->d(B_FMEAN), m->d(B_ALPHA), n, sc + 1);
    reduce_sum(s, m->d(B_LDIAG), n, sc + 2);
    int rr = check_info(m);
    if (rr) return rr;
    double h[4];
    rr = read_scalars(m, h, 4);
    if (rr) return rr;
    m->logZ = h[0] - 0.5 * h[1] - (h[2] + 0.5 * h[3]);
    m->round_corr = h[3];
    m->iters = 1;
    return GPL_OK;
}

// Laplace Newton loop for non-Gaussian likelihoods (R/laplace.R:274-297).
static int fit_laplace(gpl_model* m) {
    gpl_ctx* c = m->ctx;
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N;
    const DevProg* pd = m->b[B_PROG].as<DevProg>();
    const double* tr = m->has_trials ? m->d(B_TR) : nullptr;
    const double* off = m->has_offset ? m->d(B_OFF) : nullptr;
    {
        Prof p(c, "build_K", 0, 8.0 * N * (N + 64) / 2);
        launch_build_sym_f64(s, pd, m->prog, m->d(B_F), n, N, m->opts.jitter, 0, 1.0, nullptr, m->d(B_K), N);
    }
    fill_f64(s, m->d(B_FOLD), N, 0.0);  // fhat = 0
    const int maxiter = m->opts.maxiter > 0 ? m->opts.maxiter : 30;
    double diff = INFINITY;
    int it = 0;
    for (it = 1; it <= maxiter; ++it) {
        int* flags = m->b[B_INFO].as<int>();
        hipMemsetAsync(flags, 0, 8, s);
        lik_pseudodata_la(s, m->lik.kind, m->lik.link, lik_par(m->lik), n, m->d(B_FOLD), m->d(B_Y), tr, off, m->d(B_Z),
                          m->d(B_V), m->d(B_SV), flags + 1, (m->opts.flags & GPL_FLAG_ANALYTIC_W) ? 1 : 0);
        {
            Prof p(c, "form_B", 0, 16.0 * N * (N + 64) / 2);
            form_B_f64(s, m->d(B_K), m->d(B_L), N, n, N, m->d(B_SV));
        }
        potrf_blocked(c, m->d(B_L), N, m->d(B_DINV), m->d(B_LDIAG), flags, nullptr, nullptr, alpha_rhs(m, m->d(B_Z)));
        solve_alpha_back(c, m, m->d(B_ALPHA));
        {
            Prof p(c, "symv", 2.0 * N * N, 8.0 * N * N / 2);
            symv_lt_f64(s, m->d(B_K), N, m->d(B_ALPHA), m->d(B_FMEAN), m->d(B_KMV), kmatvec_chunks(N), 1, 0, 0);
        }
        lik_loglik(s, m->lik.kind, m->lik.link, lik_par(m->lik), n, m->d(B_FMEAN), m->d(B_Y), tr, off, m->d(B_W3));
        double* sc = m->d(B_SCAL);
        reduce_sum(s, m->d(B_W3), n, sc + 0);
        reduce_dot(s, m->d(B_FMEAN), m->d(B_ALPHA), n, sc + 1);
        reduce_sum(s, m->d(B_LDIAG), n, sc + 2);
        reduce_maxabsdiff(s, m->d(B_FOLD), m->d(B_FMEAN), n, sc + 3);
        int fl[2];
        HIP_TRY(hipMemcpyAsync(fl, flags, 8, hipMemcpyDeviceToHost, s));
        double h[4];
        int r = read_scalars(m, h, 4);
        if (r) return r;
        if (fl[1] & 2) {
            m->logZ = -INFINITY;
            return fail(GPL_ERR_NAN, "NaN in Laplace pseudo-data");
        }
        if (fl[1] & 4) {
            // some V <= 0: the split step of R/laplace.R:40-117 replaces this one
            double lz = 0;
            r = split_step(m, false, &lz);
            if (r) return r;
            m->logZ = lz;
            if (m->no_posterior) {  // fit = list(fmean = fhat_old, log_evidence = -Inf): diff 0, the loop ends
                HIP_TRY(hipMemcpyAsync(m->d(B_FMEAN), m->d(B_FOLD), N * 8, hipMemcpyDeviceToDevice, s));
                break;
            }
            reduce_maxabsdiff(s, m->d(B_FOLD), m->d(B_FMEAN), n, sc + 3);
            r = read_scalars(m, h, 4);
            if (r) return r;
        } else {
            if (fl[0] > 0) return pivot_fail(fl[0]);
            m->split = false;
            m->logZ = h[0] - 0.5 * h[1] - h[2];
        }
        diff = h[3];
        HIP_TRY(hipMemcpyAsync(m->d(B_FOLD), m->d(B_FMEAN), N * 8, hipMemcpyDeviceToDevice, s));
        if (diff < m->opts.tol) break;
    }
    if (it > maxiter) it = maxiter;
    m->iters = it;
    if (m->no_posterior) {
        m->logZ = -INFINITY;
        return fail(GPL_ERR_NAN, "the posterior of all the data is not proper (I + Lk^T V^-1 Lk is not positive "
                                 "definite, R/laplace.R:64-72): log evidence -Inf");
    }
    if (maxiter > 1 && it == maxiter) {  // also when the last allowed step converged (R/laplace.R:293)
        // the reference only warns here (R/laplace.R:293-295): keep the fit, report NOCONV
        g_err = "Maximum number of iterations in Laplace reached, max delta f = " + std::to_string(diff);
        return GPL_ERR_NOCONV;
    }
    return GPL_OK;
}

// alpha = S u with B u = S z by iterative refinement (mixed form): corrections
// from the FP32 inverse X (B^-1 ~ X^T X), residuals against the FP64 copy of B.
// Stops when the correction is below kRefineStop relative or after maxref sweeps.
static int refine_alpha_mixed(gpl_model* m, int maxref, int nch) {
    gpl_ctx* c = m->ctx;
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N;
    const float* X = m->b[B_XINV].as<float>();
    const double* B64 = m->d(B_B64);
    double* rhs = m->d(B_TERM);
    double* u = m->d(B_W1);
    double* r = m->d(B_W2);
    double* d = m->d(B_W3);
    double* t1 = m->d(B_W4);
    fill_f64(s, rhs, N, 0.0);
    vec_mul(s, n, m->d(B_SV), m->d(B_Z), rhs);
    fill_f64(s, u, N, 0.0);
    HIP_TRY(hipMemcpyAsync(r, rhs, N * 8, hipMemcpyDeviceToDevice, s));
    double rel = INFINITY, prev = INFINITY;
    int k;
    for (k = 1; k <= maxref; ++k) {
        {
            Prof p(c, "refine_solve_f32", 4.0 * N * N / 2, 4.0 * N * N);
            trmv_lower_f32(s, X, N, N, r, t1, 0, m->d(B_KMV), nch);
            trmv_lower_f32(s, X, N, N, t1, d, 1, m->d(B_KMV), nch);
        }
        vec_axpby(s, n, 1.0, d, 1.0, u, u);
        {
            Prof p(c, "symv_b64", 2.0 * N * N, 8.0 * N * (N + 64) / 2);
            symv_lt_f64(s, B64, N, u, t1, m->d(B_KMV), nch, 1, 0, 0);
        }
        vec_axpby(s, n, 1.0, rhs, -1.0, t1, r);
        rel_maxabs(s, d, u, n, m->d(B_SCAL) + 8);
        double h;
        HIP_TRY(hipMemcpyAsync(&h, m->d(B_SCAL) + 8, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        rel = h;
        if (rel < kRefineStop) break;
        if (refine_diverging(k, rel, prev)) break;
        prev = rel;
    }
    m->refine_iters_done = std::min(k, maxref);
    m->refine_rel = rel;
    if (!(rel <= kRefineTol)) return kRefineFail;
    fill_f64(s, m->d(B_ALPHA), N, 0.0);
    vec_mul(s, n, m->d(B_SV), u, m->d(B_ALPHA));
    return GPL_OK;
}

// Laplace Newton loop in the mixed-precision form: per step B = I + S K S is
// formed from the FP64 K in FP64 (kept for residuals) and FP32 (factored with
// FP64 diagonal tiles), X = L^-1 in FP32, and the Newton step's alpha =
// S B^-1 S z comes from iterative refinement against the FP64 B, so f = K alpha
// and the iterate sequence carry FP64 accuracy.  The log-determinant sums the
// FP64 pivots.  Same iteration, stopping rule and log evidence as fit_laplace.
static int fit_laplace_mixed(gpl_model* m) {
    gpl_ctx* c = m->ctx;
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N;
    const DevProg* pd = m->b[B_PROG].as<DevProg>();
    const double* tr = m->has_trials ? m->d(B_TR) : nullptr;
    const double* off = m->has_offset ? m->d(B_OFF) : nullptr;
    ENSURE(m, B_XINV, (size_t)N * N * 4);
    ENSURE(m, B_BINV, (size_t)N * N * 4);  // the inverse's scratch
    ENSURE(m, B_DACC, (size_t)N * 64 * 8);
    ENSURE(m, B_B64, (size_t)N * N * 8);
    const int nch = kmatvec_chunks(N);
    ENSURE(m, B_KMV, (size_t)N * std::max(nch, symv_slots(N, 1)) * 8);
    float* L = m->b[B_L].as<float>();
    float* X = m->b[B_XINV].as<float>();
    double* B64 = m->d(B_B64);
    {
        Prof p(c, "build_K", 0, 8.0 * N * (N + 64) / 2);
        launch_build_sym_f64(s, pd, m->prog, m->d(B_F), n, N, m->opts.jitter, 0, 1.0, nullptr, m->d(B_K), N);
    }
    fill_f64(s, m->d(B_FOLD), N, 0.0);  // fhat = 0
    const int maxiter = m->opts.maxiter > 0 ? m->opts.maxiter : 30;
    const int maxref = m->opts.refine_iters > 0 ? m->opts.refine_iters : 8;
    double diff = INFINITY;
    int it = 0;
    for (it = 1; it <= maxiter; ++it) {
        int* flags = m->b[B_INFO].as<int>();
        hipMemsetAsync(flags, 0, 8, s);
        lik_pseudodata_la(s, m->lik.kind, m->lik.link, lik_par(m->lik), n, m->d(B_FOLD), m->d(B_Y), tr, off, m->d(B_Z),
                          m->d(B_V), m->d(B_SV), flags + 1, (m->opts.flags & GPL_FLAG_ANALYTIC_W) ? 1 : 0);
        {
            Prof p(c, "form_B_mixed", 0, 20.0 * N * (N + 64) / 2);
            form_B_mixed(s, m->d(B_K), B64, L, m->d(B_DACC), N, n, N, m->d(B_SV));
        }
        potrf32_blocked(c, L, N, m->d(B_DINV), m->d(B_LDIAG), flags, m->d(B_DACC), X, m->b[B_BINV].as<float>());
        int rr0 = refine_alpha_mixed(m, maxref, nch);
        if (rr0) return rr0;
        {
            Prof p(c, "symv", 2.0 * N * N, 8.0 * N * N / 2);
            symv_lt_f64(s, m->d(B_K), N, m->d(B_ALPHA), m->d(B_FMEAN), m->d(B_KMV), nch, 1, 0, 0);
        }
        lik_loglik(s, m->lik.kind, m->lik.link, lik_par(m->lik), n, m->d(B_FMEAN), m->d(B_Y), tr, off, m->d(B_W3));
        double* sc = m->d(B_SCAL);
        reduce_sum(s, m->d(B_W3), n, sc + 0);
        reduce_dot(s, m->d(B_FMEAN), m->d(B_ALPHA), n, sc + 1);
        reduce_sum(s, m->d(B_LDIAG), n, sc + 2);
        reduce_maxabsdiff(s, m->d(B_FOLD), m->d(B_FMEAN), n, sc + 3);
        int fl[2];
        HIP_TRY(hipMemcpyAsync(fl, flags, 8, hipMemcpyDeviceToHost, s));
        double h[4];
        int rr = read_scalars(m, h, 4);
        if (rr) return rr;
        if (fl[1] & 2) {
            m->logZ = -INFINITY;
            return fail(GPL_ERR_NAN, "NaN in Laplace pseudo-data");
        }
        if (fl[1] & 4) return kRefineFail;  // some V <= 0: the split step runs in the FP64 refit
        if (fl[0] > 0) return pivot_fail(fl[0]);
        m->logZ = h[0] - 0.5 * h[1] - h[2];
        diff = h[3];
        HIP_TRY(hipMemcpyAsync(m->d(B_FOLD), m->d(B_FMEAN), N * 8, hipMemcpyDeviceToDevice, s));
        if (diff < m->opts.tol) break;
    }
    if (it > maxiter) it = maxiter;
    m->iters = it;
    m->have_xinv = true;
    if (maxiter > 1 && it == maxiter) {  // also when the last allowed step converged (R/laplace.R:293)
        g_err = "Maximum number of iterations in Laplace reached, max delta f = " + std::to_string(diff);
        return GPL_ERR_NOCONV;
    }
    return GPL_OK;
}

// probabilists' Gauss-Hermite rule: nodes = eigenvalues of the Jacobi matrix of He_k
// (Golub-Welsch), weights normalised to sum 1 -- the same rule as R/quadrature.R:28-34.
void gpl::gauss_hermite(int q, std::vector<double>& x, std::vector<double>& w) {
    // symmetric tridiagonal: diag 0, off-diagonal sqrt(k), k = 1..q-1; Jacobi eigen-iteration
    std::vector<double> A(q * q, 0.0), V(q * q, 0.0);
    for (int k = 1; k < q; ++k) A[k * q + k - 1] = A[(k - 1) * q + k] = std::sqrt((double)k);
    for (int i = 0; i < q; ++i) V[i * q + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int i = 0; i < q; ++i)
            for (int j = i + 1; j < q; ++j) off += A[i * q + j] * A[i * q + j];
        if (off < 1e-30) break;
        for (int p = 0; p < q; ++p)
            for (int r = p + 1; r < q; ++r) {
                const double apr = A[p * q + r];
                if (std::fabs(apr) < 1e-300) continue;
                const double th = (A[r * q + r] - A[p * q + p]) / (2 * apr);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
                const double cs = 1 / std::sqrt(t * t + 1), sn = t * cs;
                for (int k = 0; k < q; ++k) {
                    const double akp = A[k * q + p], akr = A[k * q + r];
                    A[k * q + p] = cs * akp - sn * akr;
                    A[k * q + r] = sn * akp + cs * akr;
                }
                for (int k = 0; k < q; ++k) {
                    const double apk = A[p * q + k], ark = A[r * q + k];
                    A[p * q + k] = cs * apk - sn * ark;
                    A[r * q + k] = sn * apk + cs * ark;
                }
                for (int k = 0; k < q; ++k) {
                    const double vkp = V[k * q + p], vkr = V[k * q + r];
                    V[k * q + p] = cs * vkp - sn * vkr;
                    V[k * q + r] = sn * vkp + cs * vkr;
                }
            }
    }
    std::vector<std::pair<double, double>> nw(q);
    for (int i = 0; i < q; ++i) nw[i] = {A[i * q + i], V[0 * q + i] * V[0 * q + i]};
    std::sort(nw.begin(), nw.end());
    x.resize(q);
    w.resize(q);
    double sw = 0;
    for (int i = 0; i < q; ++i) sw += nw[i].second;
    for (int i = 0; i < q; ++i) {
        x[i] = nw[i].first;
        w[i] = nw[i].second / sw;
    }
}

// Parallel EP (R/ep.R:286-330, ep_iter R/ep.R:16-49,140-146).
static int fit_ep(gpl_model* m) {
    gpl_ctx* c = m->ctx;
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N;
    const DevProg* pd = m->b[B_PROG].as<DevProg>();
    const double* tr = m->has_trials ? m->d(B_TR) : nullptr;
    const double* off = m->has_offset ? m->d(B_OFF) : nullptr;
    const int q = m->opts.quad_order > 0 ? m->opts.quad_order : 11;
    if (q > 32) return fail(GPL_ERR_ARG, "quad_order must be <= 32");
    std::vector<double> gx, gw;
    gauss_hermite(q, gx, gw);
    HIP_TRY(hipMemcpyAsync(m->d(B_GH), gx.data(), q * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(m->d(B_GH) + 32, gw.data(), q * 8, hipMemcpyHostToDevice, s));
    ENSURE(m, B_XINV, (size_t)N * N * 8);
    {
        Prof p(c, "build_K", 0, 8.0 * N * (N + 64) / 2);
        launch_build_sym_f64(s, pd, m->prog, m->d(B_F), n, N, m->opts.jitter, 0, 1.0, nullptr, m->d(B_K), N);
    }
    // P_old = 0, z_old = 0, fmean_old = 0, fvar_old = diag(K)
    fill_f64(s, m->d(B_POLD), N, 0.0);
    fill_f64(s, m->d(B_ZOLD), N, 0.0);
    fill_f64(s, m->d(B_FOLD), N, 0.0);
    launch_build_diag(s, pd, m->prog, m->d(B_F), N, m->d(B_F), N, n, m->d(B_FVOLD));
    // fvar_old = diag(K) with K = eval_cf + jitter I (R/gp_fit.R:154, R/ep.R:298)
    fill_f64(s, m->d(B_W3), N, m->opts.jitter);
    vec_axpby(s, n, 1.0, m->d(B_FVOLD), 1.0, m->d(B_W3), m->d(B_FVOLD));

    const int maxiter = m->opts.maxiter > 0 ? m->opts.maxiter : 100;
    const double tol = 1e-4, damping = m->opts.damping, dmin = m->opts.damping_min > 0 ? m->opts.damping_min : 0.1;
    double damp = damping, diff_old = INFINITY, diff_new = INFINITY;
    bool quad_ok = true;
    int it;
    for (it = 1; it <= maxiter; ++it) {
        int* flags = m->b[B_INFO].as<int>();
        hipMemsetAsync(flags, 0, 8, s);
        {
            Prof p(c, "ep_sites", 0, 0);
            ep_sites(s, m->lik.kind, m->lik.link, lik_par(m->lik), n, m->d(B_FOLD), m->d(B_FVOLD), m->d(B_ZOLD), m->d(B_POLD),
                     m->d(B_Y), tr, off, q, m->d(B_GH), m->d(B_GH) + 32, damp, m->d(B_Z), m->d(B_V), m->d(B_SV),
                     m->d(B_LOGZS), m->d(B_CAVM), m->d(B_CAVV), flags + 1);
        }
        int fl[2];
        HIP_TRY(hipMemcpyAsync(fl, flags, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (fl[1] & 2) {
            m->logZ = -INFINITY;
            m->iters = it;
            return fail(GPL_ERR_NAN, "NaN in EP site parameters (R/ep.R:30-32 returns log_evidence = -Inf)");
        }
        quad_ok = !(fl[1] & 1);  // the last sweep's flag decides (R/ep.R:326-328)
        double* sc = m->d(B_SCAL);
        double h[4];
        if (fl[1] & 4) {
            // some site variance <= 0: the split sweep of R/ep.R:50-138
            double lz = 0;
            int r = split_step(m, true, &lz);
            if (r) return r;
            if (m->no_posterior) {
                // The reference returns fmean = Inf (R/ep.R:90-92): diff_new = Inf, so after an
                // accepted sweep the sweep is rejected with more damping; on the first sweep it
                // would carry Inf into the next one, where it fails.
                m->logZ = -INFINITY;
                if (!(diff_old < INFINITY)) {
                    m->iters = it;
                    return fail(GPL_ERR_NAN, "EP: the posterior of all the data is not proper (I + Lk^T V^-1 Lk is "
                                             "not positive definite, R/ep.R:83-92) on the first sweep");
                }
                diff_new = INFINITY;
                damp = damp * damping;
                continue;
            }
            reduce_maxabsdiff(s, m->d(B_FOLD), m->d(B_FMEAN), n, sc + 3);
            r = read_scalars(m, h, 4);
            if (r) return r;
            m->logZ = lz;
            m->have_xinv = false;
        } else {
        {
            Prof p(c, "form_B", 0, 16.0 * N * (N + 64) / 2);
            form_B_f64(s, m->d(B_K), m->d(B_L), N, n, N, m->d(B_SV));
        }
        potrf_blocked(c, m->d(B_L), N, m->d(B_DINV), m->d(B_LDIAG), flags, nullptr, nullptr, alpha_rhs(m, m->d(B_Z)));
        solve_alpha_back(c, m, m->d(B_ALPHA));
        {
            Prof p(c, "symv", 2.0 * N * N, 8.0 * N * N / 2);
            symv_lt_f64(s, m->d(B_K), N, m->d(B_ALPHA), m->d(B_FMEAN), m->d(B_KMV), kmatvec_chunks(N), 1, 0, 0);
        }
        trtri_blocked(c, m->d(B_L), m->d(B_DINV), N, m->d(B_XINV));
        colsumsq_f64(s, m->d(B_XINV), N, N, n, m->d(B_W3), 1);
        ep_post(s, n, m->d(B_V), m->d(B_Z), m->d(B_W3), m->d(B_CAVM), m->d(B_CAVV), m->d(B_LOGZS), m->d(B_FVAR),
                m->d(B_TERM));
        reduce_sum(s, m->d(B_TERM), n, sc + 0);
        reduce_dot(s, m->d(B_Z), m->d(B_ALPHA), n, sc + 1);
        reduce_sum(s, m->d(B_LDIAG), n, sc + 2);
        reduce_maxabsdiff(s, m->d(B_FOLD), m->d(B_FMEAN), n, sc + 3);
        HIP_TRY(hipMemcpyAsync(fl, flags, 8, hipMemcpyDeviceToHost, s));
        int r = read_scalars(m, h, 4);
        if (r) return r;
        if (fl[0] > 0) return pivot_fail(fl[0]);
        m->split = false;
        m->no_posterior = false;
        m->have_xinv = true;
        m->logZ = h[0] - h[2] - 0.5 * h[1];
        }
        diff_new = h[3];
        if (diff_new > diff_old) {
            damp = damp * damping;
            continue;
        }
        damp = std::max(damp, dmin);
        HIP_TRY(hipMemcpyAsync(m->d(B_FOLD), m->d(B_FMEAN), N * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(m->d(B_FVOLD), m->d(B_FVAR), N * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(m->d(B_ZOLD), m->d(B_Z), N * 8, hipMemcpyDeviceToDevice, s));
        vec_recip(s, n, m->d(B_V), m->d(B_POLD));  // P_old = fit$P = 1 / V
        diff_old = diff_new;
        if (diff_new < tol) break;
    }
    if (it > maxiter) it = maxiter;
    m->iters = it;
    if (maxiter > 1 && it == maxiter) {  // also when the last allowed sweep converged (R/ep.R:322)
        m->logZ = -INFINITY;
        g_err = "Maximum number of iterations in EP reached, max delta f = " + std::to_string(diff_new);
        return GPL_ERR_NOCONV;
    }
    if (!quad_ok) m->logZ = -INFINITY;
    return GPL_OK;
}

// Parallel EP in the mixed-precision form: per sweep B = I + S K S is formed in
// FP64 (kept for residuals) and FP32 (factored with FP64 diagonal tiles), X = L^-1
// in FP32 gives diag(B^-1) for the posterior variances, and alpha comes from
// iterative refinement against the FP64 B.  Loop control, site updates and the
// evidence terms are those of fit_ep.
static int fit_ep_mixed(gpl_model* m) {
    gpl_ctx* c = m->ctx;
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N;
    const DevProg* pd = m->b[B_PROG].as<DevProg>();
    const double* tr = m->has_trials ? m->d(B_TR) : nullptr;
    const double* off = m->has_offset ? m->d(B_OFF) : nullptr;
    const int q = m->opts.quad_order > 0 ? m->opts.quad_order : 11;
    if (q > 32) return fail(GPL_ERR_ARG, "quad_order must be <= 32");
    std::vector<double> gx, gw;
    gauss_hermite(q, gx, gw);
    HIP_TRY(hipMemcpyAsync(m->d(B_GH), gx.data(), q * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(m->d(B_GH) + 32, gw.data(), q * 8, hipMemcpyHostToDevice, s));
    ENSURE(m, B_XINV, (size_t)N * N * 4);
    ENSURE(m, B_BINV, (size_t)N * N * 4);  // the inverse's scratch
    ENSURE(m, B_DACC, (size_t)N * 64 * 8);
    ENSURE(m, B_B64, (size_t)N * N * 8);
    const int nch = kmatvec_chunks(N);
    ENSURE(m, B_KMV, (size_t)N * std::max(nch, symv_slots(N, 1)) * 8);
    const int maxref = m->opts.refine_iters > 0 ? m->opts.refine_iters : 8;
    float* L32 = m->b[B_L].as<float>();
    float* X32 = m->b[B_XINV].as<float>();
    {
        Prof p(c, "build_K", 0, 8.0 * N * (N + 64) / 2);
        launch_build_sym_f64(s, pd, m->prog, m->d(B_F), n, N, m->opts.jitter, 0, 1.0, nullptr, m->d(B_K), N);
    }
    // P_old = 0, z_old = 0, fmean_old = 0, fvar_old = diag(K)
    fill_f64(s, m->d(B_POLD), N, 0.0);
    fill_f64(s, m->d(B_ZOLD), N, 0.0);
    fill_f64(s, m->d(B_FOLD), N, 0.0);
    launch_build_diag(s, pd, m->prog, m->d(B_F), N, m->d(B_F), N, n, m->d(B_FVOLD));
    // fvar_old = diag(K) with K = eval_cf + jitter I (R/gp_fit.R:154, R/ep.R:298)
    fill_f64(s, m->d(B_W3), N, m->opts.jitter);
    vec_axpby(s, n, 1.0, m->d(B_FVOLD), 1.0, m->d(B_W3), m->d(B_FVOLD));

    const int maxiter = m->opts.maxiter > 0 ? m->opts.maxiter : 100;
    const double tol = 1e-4, damping = m->opts.damping, dmin = m->opts.damping_min > 0 ? m->opts.damping_min : 0.1;
    double damp = damping, diff_old = INFINITY, diff_new = INFINITY;
    bool quad_ok = true;
    int it;
    for (it = 1; it <= maxiter; ++it) {
        int* flags = m->b[B_INFO].as<int>();
        hipMemsetAsync(flags, 0, 8, s);
        {
            Prof p(c, "ep_sites", 0, 0);
            ep_sites(s, m->lik.kind, m->lik.link, lik_par(m->lik), n, m->d(B_FOLD), m->d(B_FVOLD), m->d(B_ZOLD), m->d(B_POLD),
                     m->d(B_Y), tr, off, q, m->d(B_GH), m->d(B_GH) + 32, damp, m->d(B_Z), m->d(B_V), m->d(B_SV),
                     m->d(B_LOGZS), m->d(B_CAVM), m->d(B_CAVV), flags + 1);
        }
        int fl[2];
        HIP_TRY(hipMemcpyAsync(fl, flags, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (fl[1] & 2) {
            m->logZ = -INFINITY;
            m->iters = it;
            return fail(GPL_ERR_NAN, "NaN in EP site parameters (R/ep.R:30-32 returns log_evidence = -Inf)");
        }
        if (fl[1] & 4) return kRefineFail;  // some site variance <= 0: the split sweep runs in the FP64 refit
        quad_ok = !(fl[1] & 1);  // the last sweep's flag decides (R/ep.R:326-328)
        {
            Prof p(c, "form_B_mixed", 0, 20.0 * N * (N + 64) / 2);
            form_B_mixed(s, m->d(B_K), m->d(B_B64), L32, m->d(B_DACC), N, n, N, m->d(B_SV));
        }
        potrf32_blocked(c, L32, N, m->d(B_DINV), m->d(B_LDIAG), flags, m->d(B_DACC), X32, m->b[B_BINV].as<float>());
        {
            int rr0 = refine_alpha_mixed(m, maxref, nch);
            if (rr0) return rr0;
        }
        {
            Prof p(c, "symv", 2.0 * N * N, 8.0 * N * N / 2);
            symv_lt_f64(s, m->d(B_K), N, m->d(B_ALPHA), m->d(B_FMEAN), m->d(B_KMV), nch, 1, 0, 0);
        }
        colsumsq_f32(s, X32, N, N, n, m->d(B_W3), 1);
        ep_post(s, n, m->d(B_V), m->d(B_Z), m->d(B_W3), m->d(B_CAVM), m->d(B_CAVV), m->d(B_LOGZS), m->d(B_FVAR),
                m->d(B_TERM));
        double* sc = m->d(B_SCAL);
        reduce_sum(s, m->d(B_TERM), n, sc + 0);
        reduce_dot(s, m->d(B_Z), m->d(B_ALPHA), n, sc + 1);
        reduce_sum(s, m->d(B_LDIAG), n, sc + 2);
        reduce_maxabsdiff(s, m->d(B_FOLD), m->d(B_FMEAN), n, sc + 3);
        HIP_TRY(hipMemcpyAsync(fl, flags, 8, hipMemcpyDeviceToHost, s));
        double h[4];
        int r = read_scalars(m, h, 4);
        if (r) return r;
        if (fl[0] > 0) return pivot_fail(fl[0]);
        m->logZ = h[0] - h[2] - 0.5 * h[1];
        diff_new = h[3];
        if (diff_new > diff_old) {
            damp = damp * damping;
            continue;
        }
        damp = std::max(damp, dmin);
        HIP_TRY(hipMemcpyAsync(m->d(B_FOLD), m->d(B_FMEAN), N * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(m->d(B_FVOLD), m->d(B_FVAR), N * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(m->d(B_ZOLD), m->d(B_Z), N * 8, hipMemcpyDeviceToDevice, s));
        vec_recip(s, n, m->d(B_V), m->d(B_POLD));  // P_old = fit$P = 1 / V
        diff_old = diff_new;
        if (diff_new < tol) break;
    }
    if (it > maxiter) it = maxiter;
    m->iters = it;
    if (maxiter > 1 && it == maxiter) {  // also when the last allowed sweep converged (R/ep.R:322)
        m->logZ = -INFINITY;
        g_err = "Maximum number of iterations in EP reached, max delta f = " + std::to_string(diff_new);
        return GPL_ERR_NOCONV;
    }
    if (!quad_ok) m->logZ = -INFINITY;
    m->have_xinv = true;
    return GPL_OK;
}

extern "C" int gpl_fit(gpl_ctx* c, const gpl_cf_term* terms, int32_t nterms, const double* x, int64_t n,
                       int64_t ncols, const double* y, const double* trials, const double* offset, const gpl_lik* lik,
                       const gpl_opts* opts, gpl_model** inout, double* log_evidence) {
    if (!c || !x || !y || !lik || !inout) return fail(GPL_ERR_ARG, "NULL argument");
    if (n < 1 || ncols < 0) return fail(GPL_ERR_ARG, "n must be >= 1");
    std::lock_guard<std::mutex> lk(c->mu);
    Range rg("gpl_fit");
    HIP_TRY(hipSetDevice(c->device));
    gpl_opts o;
    if (opts) o = *opts;
    else gpl_opts_default(&o);
    if (lik->kind == GPL_LIK_GAUSSIAN && o.approx == GPL_APPROX_EP)
        return fail(GPL_ERR_ARG, "approx_ep disabled for lik_gaussian: use approx_laplace instead, because it gives "
                                 "exact results");  // R/gp.R:54-56
    if (lik->kind == GPL_LIK_GAUSSIAN && !(lik->sigma > 0)) return fail(GPL_ERR_ARG, "sigma must be positive");
    if (lik->kind < GPL_LIK_GAUSSIAN || lik->kind > GPL_LIK_BETABINOM) return fail(GPL_ERR_ARG, "unknown likelihood");
    if (lik->kind == GPL_LIK_BETABINOM && !(lik->phi > 0)) return fail(GPL_ERR_ARG, "phi must be positive");
    if ((lik->kind == GPL_LIK_BINOMIAL || lik->kind == GPL_LIK_BETABINOM) && !trials)
        return fail(GPL_ERR_ARG, "trials must be provided for the binomial and beta binomial likelihoods");
    if (o.precision != GPL_PREC_FP64 && o.precision != GPL_PREC_MIXED) return fail(GPL_ERR_ARG, "bad precision");
    if ((o.flags & GPL_FLAG_ANALYTIC_W) && lik->kind == GPL_LIK_BETABINOM)
        return fail(GPL_ERR_UNSUPPORTED, "the exact curvature is provided for the logit, probit and log links "
                                         "(not the beta-binomial likelihood)");
    const bool mixed = o.precision == GPL_PREC_MIXED;
    DevProg P;
    int r = prepare_prog(terms, nterms, ncols, P);
    if (r) return r;
    gpl_model* m = *inout;
    if (m && m->ctx != c) return fail(GPL_ERR_ARG, "the model to refit belongs to another context");
    bool fresh = false;
    if (!m) {
        m = new gpl_model();
        m->ctx = c;
        fresh = true;
    }
    const bool need_k = !(lik->kind == GPL_LIK_GAUSSIAN);
    r = alloc_model(m, n, ncols, need_k, mixed);
    if (r) {
        if (fresh) delete m;
        return r;
    }
    m->prog = P;
    m->lik = *lik;
    m->opts = o;
    m->has_trials = trials != nullptr;
    m->has_offset = offset != nullptr;
    m->have_xinv = m->have_binv = false;
    m->split = m->no_posterior = m->restored = false;
    hipStream_t s = c->stream;
    HIP_TRY(hipMemcpyAsync(m->b[B_PROG].p, &m->prog, sizeof(DevProg), hipMemcpyHostToDevice, s));
    if (n * ncols > 0) HIP_TRY(hipMemcpyAsync(m->d(B_X), x, n * ncols * 8, hipMemcpyHostToDevice, s));
    launch_features(s, m->d(B_X), n, n, (int)ncols, m->b[B_PROG].as<DevProg>(), m->prog, m->d(B_F), m->N);
    upload_vec(m, B_Y, y, n);
    upload_vec(m, B_TR, trials, n);
    upload_vec(m, B_OFF, offset, n);
    hipMemsetAsync(m->b[B_INFO].p, 0, 64, s);
    m->refit_fp64 = false;
    m->refine_iters_done = 0;
    m->refine_rel = 0;
    if (o.approx == GPL_APPROX_EP) r = mixed ? fit_ep_mixed(m) : fit_ep(m);
    else if (lik->kind == GPL_LIK_GAUSSIAN) r = mixed ? fit_gaussian_mixed(m) : fit_gaussian(m);
    else r = mixed ? fit_laplace_mixed(m) : fit_laplace(m);
    if (mixed && (r == kRefineFail || r == GPL_ERR_NOT_PD)) {
        // The FP32 factor could not carry this problem (refinement did not contract,
        // or rounding B to FP32 made a pivot non-positive): redo the fit in FP64,
        // whose result (or failing minor) is the one the reference would give.
        hipStreamSynchronize(s);
        for (int i : {B_B64, B_DACC, B_XINV, B_BINV}) m->b[i].free_();
        m->opts.precision = GPL_PREC_FP64;
        m->have_xinv = m->have_binv = false;
        m->round_corr = 0.0;
        m->logZ = 0.0;
        r = alloc_model(m, n, ncols, need_k, false);
        if (r) {
            if (fresh) delete m;
            else *inout = m;
            return r;
        }
        hipMemsetAsync(m->b[B_INFO].p, 0, 64, s);
        m->refit_fp64 = true;
        if (o.approx == GPL_APPROX_EP) r = fit_ep(m);
        else if (lik->kind == GPL_LIK_GAUSSIAN) r = fit_gaussian(m);
        else r = fit_laplace(m);
    }
    *inout = m;
    if (log_evidence) *log_evidence = m->logZ;
    if (r && r != GPL_ERR_NOCONV && r != GPL_ERR_NAN) return r;
    return r;
}

// d(-log evidence)/d(log theta) of a Laplace (exact curvature) or EP fit, FP64:
// -1/2 sum_ik C_ik (a_i a_k - W_ik) with W = (K + V)^-1 - (t u^T + u t^T), where
// for Laplace (R&W Algorithm 5.1) u = grad log p at the mode and
// t = (I - R K) s2, s2 = d logZ / d fhat, R = (K + V)^-1; for EP t = u = 0.
static int energy_grad_latent(gpl_model* m, double* grad) {
    gpl_ctx* c = m->ctx;
    const bool ep = m->opts.approx == GPL_APPROX_EP;
    if (m->opts.precision != GPL_PREC_FP64)
        return fail(GPL_ERR_UNSUPPORTED, "the gradient of a Laplace or EP fit is provided in FP64");
    if (m->split)
        return fail(GPL_ERR_UNSUPPORTED, "the fit split its sites (non-log-concave branch): no gradient");
    if (!ep && !(m->opts.flags & GPL_FLAG_ANALYTIC_W))
        return fail(GPL_ERR_UNSUPPORTED, "the Laplace gradient is that of the exact curvature: fit with "
                                         "GPL_FLAG_ANALYTIC_W");
    if (!ep && m->restored)
        return fail(GPL_ERR_UNSUPPORTED, "a restored model holds the fit fields but not the targets the Laplace "
                                         "gradient reads: refit it");
    std::lock_guard<std::mutex> lk(c->mu);
    Range rg("gpl_energy_grad");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N;
    ENSURE(m, B_XINV, (size_t)N * N * 8);
    ENSURE(m, B_BINV, (size_t)N * N * 8);
    const int64_t T = N / 64, ntiles = T * (T + 1) / 2;
    const int ng = m->prog.ngrad;
    ENSURE(m, B_PART, (size_t)ntiles * (ng > 0 ? ng : 1) * 8);
    ENSURE(m, B_KMV, (size_t)N * std::max(kmatvec_chunks(N), symv_slots(N, 1)) * 8);
    for (int w : {B_W1, B_W2, B_W3, B_W4}) ENSURE(m, w, (size_t)N * 8);
    if (!m->have_xinv) {
        trtri_blocked(c, m->d(B_L), m->d(B_DINV), N, m->d(B_XINV), m->d(B_BINV));
        m->have_xinv = true;
    }
    double* Bi = m->d(B_BINV);
    {
        Prof p(c, "lauum", (double)N * N * N / 3, 8.0 * N * N);
        gemm2_f64(s, 1, 0, N, N, N, 1.0, m->d(B_XINV), N, m->d(B_XINV), N, 0.0, Bi, N, 2, 0);
    }
    const double* sv = m->d(B_SV);
    double *u = nullptr, *t = nullptr;
    if (!ep) {
        u = m->d(B_W1);
        double* s2 = m->d(B_W2);
        t = m->d(B_W3);
        double* w = m->d(B_W4);
        const double* tr = m->has_trials ? m->d(B_TR) : nullptr;
        const double* off = m->has_offset ? m->d(B_OFF) : nullptr;
        if (!m->b[B_K].p) {  // a trimmed model: K is built again from the resident features
            ENSURE(m, B_K, (size_t)N * N * 8);
            launch_build_sym_f64(s, m->b[B_PROG].as<DevProg>(), m->prog, m->d(B_F), n, N, m->opts.jitter, 0, 1.0,
                                 nullptr, m->d(B_K), N);
        }
        lik_la_grad_vectors(s, m->lik.kind, m->lik.link, lik_par(m->lik), n, N, m->d(B_FMEAN), m->d(B_Y), tr, off,
                            m->d(B_V), Bi, N, u, s2);
        // t = s2 - S B^-1 S K s2 (padding of every vector stays zero)
        symv_lt_f64(s, m->d(B_K), N, s2, w, m->d(B_KMV), 0, 1, 0, 0);
        fill_f64(s, t, N, 0.0);
        vec_mul(s, n, sv, w, t);
        symv_lt_f64(s, Bi, N, t, w, m->d(B_KMV), 0, 1, 0, 0);
        vec_mul(s, n, sv, w, w);
        vec_axpby(s, n, 1.0, s2, -1.0, w, t);
    }
    form_wgrad_f64(s, Bi, N, n, sv, t, u);
    m->have_binv = false;  // B^-1 now holds the weights
    {
        Prof p(c, "grad", 0, 8.0 * N * (N + 64) / 2);
        launch_grad_tiles(s, m->b[B_PROG].as<DevProg>(), m->prog, m->d(B_F), n, N, Bi, N, 1.0, m->d(B_ALPHA),
                          m->d(B_PART));
    }
    double* sc = m->d(B_SCAL);
    reduce_cols(s, m->d(B_PART), ntiles, ng, ng, sc);
    std::vector<double> h(ng > 0 ? ng : 1);
    HIP_TRY(hipMemcpyAsync(h.data(), sc, ng * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int p = 0; p < ng; ++p) grad[p] = -0.5 * h[p];
    return GPL_OK;
}

extern "C" int gpl_energy_grad(gpl_model* m, double* grad) {
    if (!m || !grad) return fail(GPL_ERR_ARG, "NULL argument");
    if (m->lik.kind != GPL_LIK_GAUSSIAN) return energy_grad_latent(m, grad);
    gpl_ctx* c = m->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    Range rg("gpl_energy_grad");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N;
    const bool mixed = m->opts.precision == GPL_PREC_MIXED;
    const size_t esz = mixed ? 4 : 8;
    ENSURE(m, B_XINV, (size_t)N * N * esz);
    ENSURE(m, B_BINV, (size_t)N * N * esz);
    const int64_t T = N / 64, ntiles = T * (T + 1) / 2;
    const int ng = m->prog.ngrad;
    ENSURE(m, B_PART, (size_t)ntiles * (ng > 0 ? ng : 1) * 8);
    const double sigma = m->lik.sigma;
    const double s2 = sigma * sigma;
    if (mixed) {
        float* X = m->b[B_XINV].as<float>();
        float* Bi = m->b[B_BINV].as<float>();
        if (!m->have_binv) {
            Prof p(c, "lauum_f32", (double)N * N * N / 3, 4.0 * N * N);
            gemm_f32(s, 1, 0, N, N, N, 1.f, X, N, X, N, 0.f, Bi, N, 2, 0, kFoldInverse);
        }
        Prof p(c, "grad", 0, 4.0 * N * (N + 64) / 2);
        launch_grad_tiles_f32(s, m->b[B_PROG].as<DevProg>(), m->prog, m->d(B_F), n, N, Bi, N, 1.0 / s2,
                              m->d(B_ALPHA), m->d(B_PART));
    } else {
        if (!m->have_xinv) {
            trtri_blocked(c, m->d(B_L), m->d(B_DINV), N, m->d(B_XINV), m->d(B_BINV));
            m->have_xinv = true;
        }
        {
            Prof p(c, "lauum", (double)N * N * N / 3, 8.0 * N * N);
            gemm2_f64(s, 1, 0, N, N, N, 1.0, m->d(B_XINV), N, m->d(B_XINV), N, 0.0, m->d(B_BINV), N, 2, 0);
        }
        Prof p(c, "grad", 0, 8.0 * N * (N + 64) / 2);
        launch_grad_tiles(s, m->b[B_PROG].as<DevProg>(), m->prog, m->d(B_F), n, N, m->d(B_BINV), N, 1.0 / s2,
                          m->d(B_ALPHA), m->d(B_PART));
    }
    m->have_binv = true;
    double* sc = m->d(B_SCAL);
    reduce_cols(s, m->d(B_PART), ntiles, ng, ng, sc);
    reduce_dot(s, m->d(B_ALPHA), m->d(B_ALPHA), n, sc + ng);
    if (mixed) colsumsq_f32(s, m->b[B_XINV].as<float>(), N, N, n, m->d(B_W3), 1);
    else colsumsq_f64(s, m->d(B_XINV), N, N, n, m->d(B_W3), 1);
    reduce_sum(s, m->d(B_W3), n, sc + ng + 1);
    std::vector<double> h(ng + 2);
    HIP_TRY(hipMemcpyAsync(h.data(), sc, (ng + 2) * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int p = 0; p < ng; ++p) grad[p] = -0.5 * h[p];
    // sigma: -1/2 tr((a a^T - Ky^-1) 2 sigma^2 I) = -sigma^2 a^T a + tr(B^-1)
    grad[ng] = -s2 * h[ng] + h[ng + 1];
    return GPL_OK;
}

// The model's covariance program restricted to its terms idx[0..nsel), in that
// order (cfind of R/gp_pred.R:88 and R/gp_draw.R:3, summed as eval_cf.list does,
// R/cf.R:402-411).  The feature layout is unchanged, so the model's training
// features and test features made with the full program serve the subset too.
// The result is also copied to the context's device slot 9 (on stream s).
static int sub_program(gpl_ctx* c, hipStream_t s, const gpl_model* m, const int32_t* idx, int32_t nsel,
                       DevProg& P, const DevProg** dev) {
    if (!idx || nsel < 1 || nsel > GPL_MAX_TERMS) return fail(GPL_ERR_ARG, "select between 1 and 8 terms");
    P = m->prog;
    for (int k = 0; k < nsel; ++k) {
        if (idx[k] < 0 || idx[k] >= m->prog.nterms) return fail(GPL_ERR_ARG, "term index out of range");
        P.t[k] = m->prog.t[idx[k]];
    }
    P.t[0].op = GPL_OP_SUM;
    P.nterms = nsel;
    DevProg* d = (DevProg*)c->buf(9, sizeof(DevProg));
    if (!d) return fail(GPL_ERR_OOM, "device allocation failed");
    HIP_TRY(hipMemcpyAsync(d, &P, sizeof(DevProg), hipMemcpyHostToDevice, s));
    *dev = d;
    return GPL_OK;
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

size_t gpl::pred_bufs_bytes(int64_t N, int64_t ncols, int64_t M) {
    return align256((size_t)M * std::max<int64_t>(ncols, 1) * 8) + align256((size_t)M * MAX_FEAT * 8) +
           2 * align256((size_t)N * M * 8) + align256((size_t)M * 3 * 8);
}

PredBufs gpl::pred_bufs_carve(void* base, int64_t N, int64_t ncols, int64_t M) {
    char* p = (char*)base;
    PredBufs b;
    b.xt = (double*)p;
    p += align256((size_t)M * std::max<int64_t>(ncols, 1) * 8);
    b.Ft = (double*)p;
    p += align256((size_t)M * MAX_FEAT * 8);
    b.KtT = (double*)p;
    p += align256((size_t)N * M * 8);
    b.A2 = (double*)p;
    p += align256((size_t)N * M * 8);
    b.v = (double*)p;
    return b;
}

void gpl::model_dims(const gpl_model* m, int64_t* N, int64_t* ncols) {
    *N = m->N;
    *ncols = m->ncols;
}

gpl_ctx* gpl::model_ctx(const gpl_model* m) { return m->ctx; }

gpl_model* gpl::model_new(gpl_ctx* c) {
    gpl_model* m = new gpl_model();
    m->ctx = c;
    return m;
}

// Prediction on the streams of context c (which may differ from the context
// that fitted the model: the model's device buffers are read-only here).
int gpl::pred_on(gpl_ctx* c, gpl_model* m, const double* xt, int64_t mt, int32_t want_var, double* mean,
                 double* var, const int32_t* terms, int32_t nsel, const PredBufs* pb) {
    if (!c || !m || (!xt && mt > 0) || !mean || (want_var && !var)) return fail(GPL_ERR_ARG, "NULL argument");
    if (m->no_posterior) return fail(GPL_ERR_ARG, "the fit has no posterior (its log evidence is -Inf)");
    if (mt <= 0) return GPL_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    Range rg("gpl_pred");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N, M = pad64(mt);
    const DevProg* pf = m->b[B_PROG].as<DevProg>();  // features: always the full program
    const DevProg* pd = pf;
    const DevProg* hp = &m->prog;
    DevProg sub;
    if (terms) {
        int r = sub_program(c, s, m, terms, nsel, sub, &pd);
        if (r) return r;
        hp = &sub;
    }
    double* xtd = pb ? pb->xt : (double*)c->buf(0, (size_t)mt * std::max<int64_t>(m->ncols, 1) * 8);
    double* Ft = pb ? pb->Ft : (double*)c->buf(1, (size_t)M * MAX_FEAT * 8);
    double* KtT = pb ? pb->KtT : (double*)c->buf(2, (size_t)N * M * 8);
    double* vbuf = pb ? pb->v : (double*)c->buf(3, (size_t)M * 3 * 8);
    if (!xtd || !Ft || !KtT || !vbuf) return fail(GPL_ERR_OOM, "device allocation failed");
    if (mt * m->ncols > 0 && !(pb && pb->staged))
        HIP_TRY(hipMemcpyAsync(xtd, xt, mt * m->ncols * 8, hipMemcpyHostToDevice, s));
    launch_features(s, xtd, mt, mt, (int)m->ncols, pf, m->prog, Ft, M);
    {
        Prof p(c, "pred_cross", 0, 8.0 * N * M);
        launch_build_cross(s, pd, *hp, m->d(B_F), N, n, N, Ft, M, mt, M, KtT, N);
    }
    gemv_t_f64(s, KtT, N, N, mt, m->d(B_ALPHA), vbuf);
    if (want_var && m->split) {
        // var_reduction = K_t (K + V)^-1 K_t^T of the reference's C_lu branch (R/gp_pred.R:220-223)
        if (pb) return fail(GPL_ERR_UNSUPPORTED, "a fit with split sites (pseudo-variances <= 0) needs gpl_pred");
        int r = split_var_reduction(c, m, KtT, M, vbuf + M);
        if (r) return r;
        launch_build_diag(s, pd, *hp, Ft, M, Ft, M, mt, vbuf + 2 * M);
        vec_axpby(s, mt, 1.0, vbuf + 2 * M, -1.0, vbuf + M, vbuf + M);
        HIP_TRY(hipMemcpyAsync(var, vbuf + M, mt * 8, hipMemcpyDeviceToHost, s));
    } else if (want_var) {
        // aux = C^-1 K_t^T = L^-1 (sv .* K_t^T)  (R/gp_pred.R:218)
        scale_rows_f64(s, KtT, N, n, M, m->d(B_SV));
        const int64_t T = N / 64;
        if (m->have_xinv) {
            double* A2 = pb ? pb->A2 : (double*)c->buf(4, (size_t)N * M * 8);
            if (!A2) return fail(GPL_ERR_OOM, "device allocation failed");
            Prof p(c, "pred_trmm", (double)N * N * M, 8.0 * N * M * 2);
            if (m->opts.precision == GPL_PREC_MIXED)
                gemm_mixed_f32f64f64(s, 0, 0, N, M, N, 1.0, m->b[B_XINV].as<float>(), N, KtT, N, 0.0, A2, N, 3);
            else
                gemm2_f64(s, 0, 0, N, M, N, 1.0, m->d(B_XINV), N, KtT, N, 0.0, A2, N, 3, 0);
            KtT = A2;
        } else {
            for (int64_t i = 0; i < T; ++i) {
                {
                    Prof p(c, "pred_trsm_diag", 2.0 * 64 * 64 * M, 16.0 * 64 * M);
                    gemm2_f64(s, 0, 0, 64, M, 64, 1.0, m->d(B_DINV) + i * 64 * 64, 64, KtT + i * 64, N, 0.0,
                              KtT + i * 64, N, 0, 0);
                }
                if (i + 1 < T) {
                    const int64_t R = (T - 1 - i) * 64;
                    Prof p(c, "pred_trsm_update", 2.0 * R * M * 64, 16.0 * R * M);
                    gemm2_f64(s, 0, 0, R, M, 64, -1.0, m->d(B_L) + (i + 1) * 64 + i * 64 * N, N, KtT + i * 64, N,
                              1.0, KtT + (i + 1) * 64, N, 0, 0);
                }
            }
        }
        colsumsq_f64(s, KtT, N, N, mt, vbuf + M, 0);
        launch_build_diag(s, pd, *hp, Ft, M, Ft, M, mt, vbuf + 2 * M);
        vec_axpby(s, mt, 1.0, vbuf + 2 * M, -1.0, vbuf + M, vbuf + M);
        HIP_TRY(hipMemcpyAsync(var, vbuf + M, mt * 8, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipMemcpyAsync(mean, vbuf, mt * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GPL_OK;
}

// Latent posterior draws at mt test points (R/gp_draw.R:107-124, method_full,
// marginal = FALSE): mean + chol(cov + jitter I) z with the full predictive
// covariance cov = K_tt - aux^T aux, aux = C^-1 K_t^T (R/gp_pred.R:202-232).
// z (mt x draws, column-major) holds the caller's standard normal draws.
static int draw_on(gpl_ctx* c, gpl_model* m, const double* xt, int64_t mt, double jitter, const double* z,
                   int64_t draws, double* out, const int32_t* terms, int32_t nsel) {
    if (!c || !m || !xt || !z || !out || mt < 1 || draws < 1) return fail(GPL_ERR_ARG, "bad argument");
    if (m->split || m->no_posterior)
        return fail(GPL_ERR_UNSUPPORTED, "draws from a fit with split sites (pseudo-variances <= 0) are not on the "
                                         "device path");
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N, M = pad64(mt);
    const DevProg* pf = m->b[B_PROG].as<DevProg>();  // features: always the full program
    const DevProg* pd = pf;
    const DevProg* hp = &m->prog;
    DevProg sub;
    if (terms) {
        int r = sub_program(c, s, m, terms, nsel, sub, &pd);
        if (r) return r;
        hp = &sub;
    }
    double* xtd = (double*)c->buf(0, (size_t)mt * std::max<int64_t>(m->ncols, 1) * 8);
    double* Ft = (double*)c->buf(1, (size_t)M * MAX_FEAT * 8);
    double* KtT = (double*)c->buf(2, (size_t)N * M * 8);
    double* vbuf = (double*)c->buf(3, (size_t)M * 3 * 8);
    double* A2 = (double*)c->buf(4, (size_t)N * M * 8);
    double* Cv = (double*)c->buf(5, (size_t)M * M * 8);
    double* Dv = (double*)c->buf(6, (size_t)M * 64 * 8 + (size_t)M * 8 + 64);
    const int64_t Dp = pad64(draws);  // the GEMM writes whole 16-column blocks
    double* Zd = (double*)c->buf(7, (size_t)M * Dp * 8);
    double* Wd = (double*)c->buf(8, (size_t)M * Dp * 8);
    if (!xtd || !Ft || !KtT || !vbuf || !A2 || !Cv || !Dv || !Zd || !Wd)
        return fail(GPL_ERR_OOM, "device allocation failed");
    double* ld = Dv + M * 64;
    int* info = (int*)(ld + M);
    if (mt * m->ncols > 0) HIP_TRY(hipMemcpyAsync(xtd, xt, mt * m->ncols * 8, hipMemcpyHostToDevice, s));
    launch_features(s, xtd, mt, mt, (int)m->ncols, pf, m->prog, Ft, M);
    launch_build_cross(s, pd, *hp, m->d(B_F), N, n, N, Ft, M, mt, M, KtT, N);
    gemv_t_f64(s, KtT, N, N, mt, m->d(B_ALPHA), vbuf);
    scale_rows_f64(s, KtT, N, n, M, m->d(B_SV));
    const int64_t T = N / 64;
    if (m->have_xinv) {
        if (m->opts.precision == GPL_PREC_MIXED)
            gemm_mixed_f32f64f64(s, 0, 0, N, M, N, 1.0, m->b[B_XINV].as<float>(), N, KtT, N, 0.0, A2, N, 3);
        else
            gemm2_f64(s, 0, 0, N, M, N, 1.0, m->d(B_XINV), N, KtT, N, 0.0, A2, N, 3, 0);
    } else {
        for (int64_t i = 0; i < T; ++i) {
            gemm2_f64(s, 0, 0, 64, M, 64, 1.0, m->d(B_DINV) + i * 64 * 64, 64, KtT + i * 64, N, 0.0, KtT + i * 64, N,
                      0, 0);
            if (i + 1 < T)
                gemm2_f64(s, 0, 0, (T - 1 - i) * 64, M, 64, -1.0, m->d(B_L) + (i + 1) * 64 + i * 64 * N, N,
                          KtT + i * 64, N, 1.0, KtT + (i + 1) * 64, N, 0, 0);
        }
        A2 = KtT;
    }
    // cov = K_tt + jitter I - aux^T aux over lower tiles (padding: identity), then its factor
    launch_build_sym_f64(s, pd, *hp, Ft, mt, M, jitter, 0, 1.0, nullptr, Cv, M);
    {
        Prof p(c, "draw_cov", (double)M * M * N, 8.0 * N * M);
        gemm2_f64(s, 1, 0, M, M, N, -1.0, A2, N, A2, N, 1.0, Cv, M, 1, 0);
    }
    HIP_TRY(hipMemsetAsync(info, 0, 64, s));
    potrf_blocked(c, Cv, M, Dv, ld, info);
    int inf = 0;
    HIP_TRY(hipMemcpyAsync(&inf, info, 4, hipMemcpyDeviceToHost, s));
    // W = L Z (Z padded with zero rows); out = mean + W
    fill_f64(s, Zd, M * Dp, 0.0);
    HIP_TRY(upload_2d(c, s, Zd, M, z, mt, draws, 20));
    {
        Prof p(c, "draw_trmm", (double)M * M * Dp, 8.0 * M * Dp * 2);
        gemm2_f64(s, 0, 0, M, Dp, M, 1.0, Cv, M, Zd, M, 0.0, Wd, M, 3, 0);
    }
    std::vector<double> mean(mt);
    HIP_TRY(hipMemcpyAsync(mean.data(), vbuf, mt * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(download_2d(c, s, out, Wd, M, mt, draws, 21));
    HIP_TRY(hipStreamSynchronize(s));
    if (inf > 0) {
        set_minor(inf);
        return fail(GPL_ERR_NOT_PD, "predictive covariance: the leading minor of order " + std::to_string(inf) +
                                        " is not positive (raise jitter)");
    }
    for (int64_t j = 0; j < draws; ++j)
        for (int64_t i = 0; i < mt; ++i) out[i + j * mt] += mean[i];
    return GPL_OK;
}

// Draws from the prior (gp_draw_prior.method_full, R/gp_draw.R:56-67): chol(K_tt + jitter I) z
// (gp_pred_prior with cov = TRUE, R/gp_pred.R:142-151; the prior mean is 0).
extern "C" int gpl_draw_prior(gpl_ctx* c, const gpl_cf_term* terms, int32_t nterms, const double* xt, int64_t mt,
                              int64_t ncols, double jitter, const double* z, int64_t draws, double* out) {
    if (!c || !xt || !z || !out || mt < 1 || draws < 1 || ncols < 0) return fail(GPL_ERR_ARG, "bad argument");
    DevProg P;
    int r = prepare_prog(terms, nterms, ncols, P);
    if (r) return r;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t M = pad64(mt), Dp = pad64(draws);
    DevProg* pd = (DevProg*)c->buf(9, sizeof(DevProg));
    double* xtd = (double*)c->buf(0, (size_t)mt * std::max<int64_t>(ncols, 1) * 8);
    double* Ft = (double*)c->buf(1, (size_t)M * MAX_FEAT * 8);
    double* Cv = (double*)c->buf(5, (size_t)M * M * 8);
    double* Dv = (double*)c->buf(6, (size_t)M * 64 * 8 + (size_t)M * 8 + 64);
    double* Zd = (double*)c->buf(7, (size_t)M * Dp * 8);
    double* Wd = (double*)c->buf(8, (size_t)M * Dp * 8);
    if (!pd || !xtd || !Ft || !Cv || !Dv || !Zd || !Wd) return fail(GPL_ERR_OOM, "device allocation failed");
    double* ld = Dv + M * 64;
    int* info = (int*)(ld + M);
    HIP_TRY(hipMemcpyAsync(pd, &P, sizeof(DevProg), hipMemcpyHostToDevice, s));
    if (mt * ncols > 0) HIP_TRY(hipMemcpyAsync(xtd, xt, mt * ncols * 8, hipMemcpyHostToDevice, s));
    launch_features(s, xtd, mt, mt, (int)ncols, pd, P, Ft, M);
    launch_build_sym_f64(s, pd, P, Ft, mt, M, jitter, 0, 1.0, nullptr, Cv, M);
    HIP_TRY(hipMemsetAsync(info, 0, 64, s));
    potrf_blocked(c, Cv, M, Dv, ld, info);
    int inf = 0;
    HIP_TRY(hipMemcpyAsync(&inf, info, 4, hipMemcpyDeviceToHost, s));
    fill_f64(s, Zd, M * Dp, 0.0);
    HIP_TRY(upload_2d(c, s, Zd, M, z, mt, draws, 20));
    gemm2_f64(s, 0, 0, M, Dp, M, 1.0, Cv, M, Zd, M, 0.0, Wd, M, 3, 0);
    HIP_TRY(download_2d(c, s, out, Wd, M, mt, draws, 21));
    HIP_TRY(hipStreamSynchronize(s));
    if (inf > 0) {
        set_minor(inf);
        return fail(GPL_ERR_NOT_PD, "prior covariance: the leading minor of order " + std::to_string(inf) +
                                        " is not positive (raise jitter)");
    }
    return GPL_OK;
}

extern "C" int gpl_draw(gpl_model* m, const double* xt, int64_t mt, double jitter, const double* z, int64_t draws,
                        double* out) {
    if (!m) return fail(GPL_ERR_ARG, "NULL model");
    return draw_on(m->ctx, m, xt, mt, jitter, z, draws, out, nullptr, 0);
}

extern "C" int gpl_draw_terms(gpl_model* m, const int32_t* terms, int32_t nsel, const double* xt, int64_t mt,
                              double jitter, const double* z, int64_t draws, double* out) {
    if (!m || !terms) return fail(GPL_ERR_ARG, "NULL argument");
    return draw_on(m->ctx, m, xt, mt, jitter, z, draws, out, terms, nsel);
}

extern "C" int gpl_pred_terms(gpl_model* m, const int32_t* terms, int32_t nsel, const double* xt, int64_t mt,
                              int32_t want_var, double* mean, double* var) {
    if (!m || !terms) return fail(GPL_ERR_ARG, "NULL argument");
    return pred_on(m->ctx, m, xt, mt, want_var, mean, var, terms, nsel);
}

extern "C" int gpl_pred(gpl_model* m, const double* xt, int64_t mt, int32_t want_var, double* mean, double* var) {
    if (!m) return fail(GPL_ERR_ARG, "NULL model");
    return pred_on(m->ctx, m, xt, mt, want_var, mean, var);
}

extern "C" int gpl_pred_ctx(gpl_ctx* c, gpl_model* m, const double* xt, int64_t mt, int32_t want_var, double* mean,
                            double* var) {
    return pred_on(c, m, xt, mt, want_var, mean, var);
}

extern "C" int64_t gpl_model_bytes(gpl_model* m) {
    if (!m) return 0;
    int64_t b = 0;
    for (auto& x : m->b) b += (int64_t)x.sz;
    return b;
}

extern "C" int gpl_model_vector(gpl_model* m, int32_t which, double* out) {
    if (!m || !out) return fail(GPL_ERR_ARG, "NULL argument");
    int idx;
    switch (which) {
        case GPL_VEC_FMEAN: idx = B_FMEAN; break;
        case GPL_VEC_ALPHA: idx = B_ALPHA; break;
        case GPL_VEC_Z: idx = B_Z; break;
        case GPL_VEC_V: idx = B_V; break;
        case GPL_VEC_FVAR: idx = B_FVAR; break;
        case GPL_VEC_CAVITY_MEAN: idx = B_CAVM; break;
        case GPL_VEC_CAVITY_VAR: idx = B_CAVV; break;
        default: return fail(GPL_ERR_ARG, "unknown vector");
    }
    std::lock_guard<std::mutex> lk(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    HIP_TRY(hipMemcpyAsync(out, m->d(idx), m->n * 8, hipMemcpyDeviceToHost, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return GPL_OK;
}

extern "C" int gpl_model_chol(gpl_model* m, double* out) {
    if (!m || !out) return fail(GPL_ERR_ARG, "NULL argument");
    if (m->split || m->no_posterior)
        return fail(GPL_ERR_UNSUPPORTED, "a fit with split sites holds no C_chol (the reference keeps C_lu)");
    std::lock_guard<std::mutex> lk(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    const int64_t n = m->n, N = m->N;
    std::vector<double> V(n);
    if (m->opts.precision == GPL_PREC_MIXED) {  // FP32 factor
        std::vector<float> f((size_t)n * n);
        HIP_TRY(hipMemcpy2DAsync(f.data(), n * 4, m->b[B_L].p, N * 4, n * 4, n, hipMemcpyDeviceToHost,
                                 m->ctx->stream));
        HIP_TRY(hipStreamSynchronize(m->ctx->stream));
        for (size_t e = 0; e < f.size(); ++e) out[e] = (double)f[e];
    } else {
        HIP_TRY(hipMemcpy2DAsync(out, n * 8, m->d(B_L), N * 8, n * 8, n, hipMemcpyDeviceToHost, m->ctx->stream));
    }
    HIP_TRY(hipMemcpyAsync(V.data(), m->d(B_V), n * 8, hipMemcpyDeviceToHost, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) out[i + j * n] = (i >= j) ? std::sqrt(V[i]) * out[i + j * n] : 0.0;
    return GPL_OK;
}

// gp_load (R/saveload.R:37-41): a device model from the saved gp$fit fields,
// without refitting.  The saved C_chol goes to the factor buffer and is
// rescaled in place; the diagonal-block inverses are recomputed from it.
extern "C" int gpl_model_restore(gpl_ctx* c, const gpl_cf_term* terms, int32_t nterms, const double* x, int64_t n,
                                 int64_t ncols, const gpl_lik* lik, const gpl_opts* opts, const double* c_chol,
                                 const double* const* vectors, double log_evidence, int32_t iters, gpl_model** out) {
    if (!c || !x || !lik || !c_chol || !vectors || !out) return fail(GPL_ERR_ARG, "NULL argument");
    if (n < 1 || ncols < 0) return fail(GPL_ERR_ARG, "n must be >= 1");
    const double* V = vectors[GPL_VEC_V];
    if (!V || !vectors[GPL_VEC_ALPHA]) return fail(GPL_ERR_ARG, "alpha and V are required");
    for (int64_t i = 0; i < n; ++i)
        if (!(V[i] > 0.0) || !std::isfinite(V[i])) return fail(GPL_ERR_ARG, "V must be positive and finite");
    gpl_opts o;
    if (opts) o = *opts;
    else gpl_opts_default(&o);
    if (lik->kind == GPL_LIK_GAUSSIAN && o.approx == GPL_APPROX_EP)
        return fail(GPL_ERR_ARG, "approx_ep disabled for lik_gaussian");
    o.precision = GPL_PREC_FP64;  // the saved factor is FP64
    DevProg P;
    int r = prepare_prog(terms, nterms, ncols, P);
    if (r) return r;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    gpl_model* m = new gpl_model();
    m->ctx = c;
    r = alloc_model(m, n, ncols, false, false);
    if (r) {
        delete m;
        return r;
    }
    m->prog = P;
    m->lik = *lik;
    m->opts = o;
    m->logZ = log_evidence;
    m->iters = iters;
    m->restored = true;
    const int64_t N = m->N;
    hipStream_t s = c->stream;
    int info = 0;
    auto upload = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(m->b[B_PROG].p, &m->prog, sizeof(DevProg), hipMemcpyHostToDevice, s));
        if (n * ncols > 0) HIP_TRY(hipMemcpyAsync(m->d(B_X), x, n * ncols * 8, hipMemcpyHostToDevice, s));
        launch_features(s, m->d(B_X), n, n, (int)ncols, m->b[B_PROG].as<DevProg>(), m->prog, m->d(B_F), N);
        for (int idx : {B_Y, B_TR, B_OFF, B_SV}) upload_vec(m, idx, nullptr, n);
        const int map[7] = {B_FMEAN, B_ALPHA, B_Z, B_V, B_FVAR, B_CAVM, B_CAVV};
        for (int w = 0; w < 7; ++w) upload_vec(m, map[w], vectors[w], n);
        HIP_TRY(hipMemcpy2DAsync(m->d(B_L), N * 8, c_chol, n * 8, n * 8, n, hipMemcpyHostToDevice, s));
        restore_factor_f64(s, m->d(B_L), n, N, m->d(B_V), m->d(B_SV));
        HIP_TRY(hipMemsetAsync(m->b[B_INFO].p, 0, 64, s));
        trtri_diag64_f64(s, m->d(B_L), N, N / 64, m->d(B_DINV), m->d(B_LDIAG), m->b[B_INFO].as<int>());
        HIP_TRY(hipMemcpyAsync(&info, m->b[B_INFO].p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        return GPL_OK;
    };
    r = upload();
    if (r) {
        delete m;
        return r;
    }
    if (info > 0) {
        delete m;
        set_minor(info);
        return fail(GPL_ERR_NOT_PD, "saved C_chol: the diagonal entry " + std::to_string(info) + " is not positive");
    }
    *out = m;
    return GPL_OK;
}

extern "C" int gpl_model_precision(gpl_model* m, int32_t* precision, int32_t* refit_fp64, int32_t* refine_sweeps,
                                   double* refine_rel) {
    if (!m) return fail(GPL_ERR_ARG, "NULL model");
    if (precision) *precision = m->opts.precision;
    if (refit_fp64) *refit_fp64 = m->refit_fp64 ? 1 : 0;
    if (refine_sweeps) *refine_sweeps = m->refine_iters_done;
    if (refine_rel) *refine_rel = m->refine_rel;
    return GPL_OK;
}

// Frees what only the fit needed (K, the FP64 copy of B, B^-1, workspaces); the
// model keeps what prediction, draws and gp_save read.  A later gradient or
// refit allocates again what it needs.
extern "C" int gpl_model_trim(gpl_model* m) {
    if (!m) return fail(GPL_ERR_ARG, "NULL model");
    std::lock_guard<std::mutex> lk(m->ctx->mu);
    HIP_TRY(hipSetDevice(m->ctx->device));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    for (int i : {B_K, B_B64, B_BINV, B_DACC, B_PART, B_SPG, B_SPH, B_SPQ}) m->b[i].free_();
    m->have_binv = false;
    return GPL_OK;
}

extern "C" int gpl_model_info(gpl_model* m, int64_t* n, int32_t* iters, double* log_evidence) {
    if (!m) return fail(GPL_ERR_ARG, "NULL model");
    if (n) *n = m->n;
    if (iters) *iters = m->iters;
    if (log_evidence) *log_evidence = m->logZ;
    return GPL_OK;
}

extern "C" void gpl_model_free(gpl_model* m) {
    if (!m) return;
    hipSetDevice(m->ctx->device);
    hipStreamSynchronize(m->ctx->stream);
    delete m;
}

// ---------------------------------------------------------------------------
// Level 0: eval_cf between two point sets (host in, host out)
// ---------------------------------------------------------------------------
extern "C" int gpl_eval_cf(gpl_ctx* c, const gpl_cf_term* terms, int32_t nterms, const double* x1, int64_t n1,
                           const double* x2, int64_t n2, int64_t ncols, int32_t diag_only, double* out) {
    if (!c || !out || (!x1 && n1 > 0) || (!x2 && n2 > 0)) return fail(GPL_ERR_ARG, "NULL argument");
    if (n1 < 0 || n2 < 0) return fail(GPL_ERR_ARG, "negative size");
    DevProg P;
    int r = prepare_prog(terms, nterms, ncols, P);
    if (r) return r;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t N1 = pad64(std::max<int64_t>(n1, 1)), N2 = pad64(std::max<int64_t>(n2, 1));
    const int64_t nc = std::max<int64_t>(ncols, 1);
    DevProg* pd = (DevProg*)c->buf(10, sizeof(DevProg));
    double* X1 = (double*)c->buf(11, (size_t)std::max<int64_t>(n1, 1) * nc * 8);
    double* X2 = (double*)c->buf(12, (size_t)std::max<int64_t>(n2, 1) * nc * 8);
    double* F1 = (double*)c->buf(13, (size_t)N1 * MAX_FEAT * 8);
    double* F2 = (double*)c->buf(14, (size_t)N2 * MAX_FEAT * 8);
    if (!pd || !X1 || !X2 || !F1 || !F2) return fail(GPL_ERR_OOM, "device allocation failed");
    HIP_TRY(hipMemcpyAsync(pd, &P, sizeof(DevProg), hipMemcpyHostToDevice, s));
    if (n1 * ncols > 0) HIP_TRY(hipMemcpyAsync(X1, x1, n1 * ncols * 8, hipMemcpyHostToDevice, s));
    if (n2 * ncols > 0) HIP_TRY(hipMemcpyAsync(X2, x2, n2 * ncols * 8, hipMemcpyHostToDevice, s));
    launch_features(s, X1, n1, n1, (int)ncols, pd, P, F1, N1);
    launch_features(s, X2, n2, n2, (int)ncols, pd, P, F2, N2);
    if (diag_only) {
        const int64_t nd = std::min(n1, n2);
        double* o = (double*)c->buf(15, (size_t)std::max<int64_t>(nd, 1) * 8);
        if (!o) return fail(GPL_ERR_OOM, "device allocation failed");
        launch_build_diag(s, pd, P, F1, N1, F2, N2, nd, o);
        if (nd > 0) HIP_TRY(hipMemcpyAsync(out, o, nd * 8, hipMemcpyDeviceToHost, s));
    } else {
        double* o = (double*)c->buf(15, (size_t)N1 * N2 * 8);
        if (!o) return fail(GPL_ERR_OOM, "device allocation failed");
        {
            Prof p(c, "build_cross", 0, 8.0 * N1 * N2);
            launch_build_cross(s, pd, P, F1, N1, n1, N1, F2, N2, n2, N2, o, N1);
        }
        if (n1 > 0 && n2 > 0)
            HIP_TRY(hipMemcpy2DAsync(out, n1 * 8, o, N1 * 8, n1 * 8, n2, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return GPL_OK;
}

static int cf_simple(gpl_ctx* c, int kind, const double* x1, int64_t n1, const double* x2, int64_t n2, int64_t d,
                     const double* lscale, int64_t nl, double magn, double sigma0, double sigma, int32_t diag_only,
                     double* out) {
    if (d < 1 || d > GPL_MAX_COLS) return fail(GPL_ERR_ARG, "d must be in [1, 16]");
    gpl_cf_term t;
    memset(&t, 0, sizeof(t));
    t.kind = kind;
    t.op = GPL_OP_SUM;
    t.col0 = 0;
    t.ncols = (int32_t)d;
    t.magn = magn;
    t.sigma0 = sigma0;
    t.sigma = sigma;
    if (kind != GPL_CF_NN) {
        if (!lscale || (nl != 1 && nl != d)) return fail(GPL_ERR_ARG, "lscale must have length 1 or d");
        t.n_lscale = (int32_t)nl;
        for (int64_t k = 0; k < nl; ++k) t.lscale[k] = lscale[k];
    }
    return gpl_eval_cf(c, &t, 1, x1, n1, x2, n2, d, diag_only, out);
}

extern "C" int gpl_cf_sexp(gpl_ctx* c, const double* x1, int64_t n1, const double* x2, int64_t n2, int64_t d,
                           const double* lscale, int64_t n_lscale, double magn, int32_t diag_only, double* out) {
    return cf_simple(c, GPL_CF_SEXP, x1, n1, x2, n2, d, lscale, n_lscale, magn, 0, 0, diag_only, out);
}
extern "C" int gpl_cf_matern32(gpl_ctx* c, const double* x1, int64_t n1, const double* x2, int64_t n2, int64_t d,
                               const double* lscale, int64_t n_lscale, double magn, int32_t diag_only, double* out) {
    return cf_simple(c, GPL_CF_MATERN32, x1, n1, x2, n2, d, lscale, n_lscale, magn, 0, 0, diag_only, out);
}
extern "C" int gpl_cf_matern52(gpl_ctx* c, const double* x1, int64_t n1, const double* x2, int64_t n2, int64_t d,
                               const double* lscale, int64_t n_lscale, double magn, int32_t diag_only, double* out) {
    return cf_simple(c, GPL_CF_MATERN52, x1, n1, x2, n2, d, lscale, n_lscale, magn, 0, 0, diag_only, out);
}
extern "C" int gpl_cf_nn(gpl_ctx* c, const double* x1, int64_t n1, const double* x2, int64_t n2, int64_t d,
                         double sigma0, double sigma, double magn, int32_t diag_only, double* out) {
    return cf_simple(c, GPL_CF_NN, x1, n1, x2, n2, d, nullptr, 1, magn, sigma0, sigma, diag_only, out);
}
