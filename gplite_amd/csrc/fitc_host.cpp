// FITC (method_fitc) on the device: the n x m cross covariance K_xz stays
// resident in both orientations, and every Newton step reduces to one
// K_zx S^-1 K_xz product on MFMA, an m x m Cholesky and matrix-vector work
// (kernels of its own: fitc.hip).
//
// Reference: R/gp_fit.R:101-123 (fit_laplace.method_fitc),
// R/laplace.R:122-180 (laplace_iter.method_fitc, log-concave branch),
// R/laplace.R:300-322 (laplace.method_fitc), R/linear_algebra.R:3-172
// (inv_lemma_get / inv_lemma_solve), R/gp_pred.R:234-290
// (gp_pred_post.method_fitc).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gplite_amd.h"
#include "host_common.h"

using namespace gpl;

namespace {

enum {
    F_PROG, F_X, F_F, F_ZX, F_FZ, F_Y, F_TR, F_OFF,
    F_KZ, F_LZ, F_DINVZ, F_LDZ, F_XZ,           // Kz (kept), its factor, diagonal inverses, log pivots, Lz^-1
    F_U, F_UT, F_W,                              // K_xz (N x Mz), K_zx (Mz x N), S^-1/2 K_xz
    F_H, F_G, F_DINVG, F_LDG, F_XG, F_MM, F_T1,  // H = K_zx S^-1 K_xz, G = Kz + H -> Lc, Lc^-1, M, scratch
    F_WZ,                                        // w = Lz^-1 K_zx alpha
    F_D, F_Z, F_V, F_SV, F_S, F_RS, F_ZS, F_ALPHA, F_FMEAN, F_FOLD, F_V1, F_V2, F_V3, F_V4,
    F_FVAR, F_FVOLD, F_ZOLD, F_POLD, F_CAVM, F_CAVV, F_LOGZS, F_TERM, F_GH,  // EP state
    F_Q1, F_Q2, F_SCAL, F_INFO, F_SPLIT,
    F_COUNT
};

}  // namespace

struct gpl_fitc {
    gpl_ctx* ctx = nullptr;
    int64_t n = 0, N = 0, m = 0, Mz = 0, ncols = 0;
    DevProg prog{};
    gpl_lik lik{};
    gpl_opts opts{};
    bool has_trials = false, has_offset = false;
    double logZ = 0;
    int32_t iters = 0;
    DevBuf b[F_COUNT];
    ~gpl_fitc() {
        for (auto& x : b) x.free_();
    }
    double* d(int i) const { return b[i].as<double>(); }
};

#define FENSURE(fm, idx, bytes)                                                               \
    do {                                                                                      \
        if (!(fm)->b[idx].ensure(bytes)) return fail(GPL_ERR_OOM, "device allocation failed"); \
    } while (0)

// K-split of the K_zx S^-1 K_xz product: with few m x m output tiles and a long
// reduction over n, P partial products (rows split in P ranges) fill the chip and
// are summed in a fixed order.
static int fitc_split(const gpl_fitc* f) {
    const int64_t tiles = (f->Mz / 64) * (f->Mz / 64);
    int P = 1;
    while (P < 16 && tiles * P < 1024 && (f->N / 64) % (2 * P) == 0 && f->N / (2 * P) >= 512) P *= 2;
    return P;
}

static int fitc_alloc(gpl_fitc* f) {
    const int64_t N = f->N, Mz = f->Mz;
    const size_t vN = (size_t)N * 8, vM = (size_t)Mz * 8, mm = (size_t)Mz * Mz * 8, nm = (size_t)N * Mz * 8;
    FENSURE(f, F_PROG, sizeof(DevProg));
    FENSURE(f, F_X, (size_t)f->n * std::max<int64_t>(f->ncols, 1) * 8);
    FENSURE(f, F_ZX, (size_t)f->m * std::max<int64_t>(f->ncols, 1) * 8);
    FENSURE(f, F_F, (size_t)N * MAX_FEAT * 8);
    FENSURE(f, F_FZ, (size_t)Mz * MAX_FEAT * 8);
    for (int i : {F_KZ, F_LZ, F_XZ, F_H, F_G, F_XG, F_MM, F_T1}) FENSURE(f, i, mm);
    for (int i : {F_DINVZ, F_DINVG}) FENSURE(f, i, (size_t)Mz * 64 * 8);
    for (int i : {F_LDZ, F_LDG, F_WZ, F_Q1, F_Q2}) FENSURE(f, i, vM);
    for (int i : {F_U, F_UT, F_W}) FENSURE(f, i, nm);
    for (int i : {F_Y, F_TR, F_OFF, F_D, F_Z, F_V, F_SV, F_S, F_RS, F_ZS, F_ALPHA, F_FMEAN, F_FOLD, F_V1, F_V2, F_V3,
                  F_V4, F_FVAR, F_FVOLD, F_ZOLD, F_POLD, F_CAVM, F_CAVV, F_LOGZS, F_TERM})
        FENSURE(f, i, vN);
    FENSURE(f, F_GH, 64 * 8);
    FENSURE(f, F_SCAL, 64 * 8);
    FENSURE(f, F_INFO, 64);
    const int P = fitc_split(f);
    if (P > 1) FENSURE(f, F_SPLIT, (size_t)P * Mz * Mz * 8);
    return GPL_OK;
}

static int fitc_upload(gpl_fitc* f, int idx, const double* h) {
    hipStream_t s = f->ctx->stream;
    fill_f64(s, f->d(idx), f->N, 0.0);
    if (h) HIP_TRY(hipMemcpyAsync(f->d(idx), h, f->n * 8, hipMemcpyHostToDevice, s));
    return GPL_OK;
}

static int fitc_scalars(gpl_fitc* f, double* out, int k, int* flags) {
    hipStream_t s = f->ctx->stream;
    HIP_TRY(hipMemcpyAsync(out, f->d(F_SCAL), k * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(flags, f->b[F_INFO].p, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GPL_OK;
}

// M = A C^-1 A^T with A = Lz^-1 K_zx, the m x m matrix of the predictive
// variance (inv_lemma_solve(C_inv, t(Kt_aux1), Kt_aux1), R/gp_pred.R:279):
// A S^-1 A^T - (A S^-1 K_xz) G^-1 (K_zx S^-1 A^T) = Lz^-1 (H - H G^-1 H) Lz^-T.
static void fitc_predictive_matrix(gpl_fitc* f) {
    gpl_ctx* c = f->ctx;
    hipStream_t s = c->stream;
    const int64_t Mz = f->Mz;
    trtri_blocked(c, f->d(F_G), f->d(F_DINVG), Mz, f->d(F_XG));
    double* Y = f->d(F_MM);  // Y = Lc^-1 H, then reused for M
    gemm2_f64(s, 0, 0, Mz, Mz, Mz, 1.0, f->d(F_XG), Mz, f->d(F_H), Mz, 0.0, Y, Mz, 3, 0);
    hipMemcpyAsync(f->d(F_T1), f->d(F_H), Mz * Mz * 8, hipMemcpyDeviceToDevice, s);
    gemm2_f64(s, 1, 0, Mz, Mz, Mz, -1.0, Y, Mz, Y, Mz, 1.0, f->d(F_T1), Mz, 0, 0);  // H - H G^-1 H
    gemm2_f64(s, 0, 0, Mz, Mz, Mz, 1.0, f->d(F_XZ), Mz, f->d(F_T1), Mz, 0.0, f->d(F_XG), Mz, 3, 0);
    gemm2_f64(s, 0, 1, Mz, Mz, Mz, 1.0, f->d(F_XG), Mz, f->d(F_XZ), Mz, 0.0, Y, Mz, 0, 0);
}

// One inversion-lemma solve for the current sites: S = D + V, G = Kz + K_zx S^-1 K_xz
// and its factor, alpha = C^-1 z and the posterior mean K_f alpha into F_FMEAN
// (R/laplace.R:134-141, R/ep.R:166-176).  flags[0] receives G's failing minor.
static void fitc_lemma_step(gpl_fitc* f, int* flags) {
    gpl_ctx* c = f->ctx;
    hipStream_t s = c->stream;
    const int64_t n = f->n, N = f->N, Mz = f->Mz;
    fitc_site(s, n, N, f->d(F_D), f->d(F_V), f->d(F_Z), f->d(F_S), f->d(F_RS), f->d(F_ZS));
    scale_rows_copy_f64(s, f->d(F_U), N, N, Mz, f->d(F_RS), f->d(F_W));
    {
        Prof p(c, "fitc_syrk", 2.0 * Mz * Mz * N, 8.0 * N * Mz);
        const int P = fitc_split(f);
        if (P == 1) {
            gemm2_f64(s, 1, 0, Mz, Mz, N, 1.0, f->d(F_W), N, f->d(F_W), N, 0.0, f->d(F_H), Mz, 0, 0);
        } else {
            const int64_t Kp = N / P;
            gemm2_f64_batched(s, 1, 0, Mz, Mz, Kp, 1.0, f->d(F_W), N, f->d(F_W), N, 0.0, f->d(F_SPLIT), Mz, 0, P, Kp,
                              Kp, Mz * Mz);
            sum_parts_f64(s, f->d(F_SPLIT), Mz * Mz, P, Mz * Mz, f->d(F_H));
        }
    }
    vec_axpby(s, Mz * Mz, 1.0, f->d(F_KZ), 1.0, f->d(F_H), f->d(F_G));
    potrf_blocked(c, f->d(F_G), Mz, f->d(F_DINVG), f->d(F_LDG), flags);
    Prof pv(c, "fitc_gemv", 0, 32.0 * N * Mz);
    gemv_t_f64(s, f->d(F_U), N, N, Mz, f->d(F_ZS), f->d(F_Q1));
    trsv_lower(c, s, f->d(F_G), Mz, f->d(F_DINVG), Mz, f->d(F_Q1), f->d(F_Q2), 0, flags);
    trsv_lower(c, s, f->d(F_G), Mz, f->d(F_DINVG), Mz, f->d(F_Q2), f->d(F_Q1), 1, flags);
    gemv_t_f64(s, f->d(F_UT), Mz, Mz, N, f->d(F_Q1), f->d(F_V1));
    fitc_alpha(s, n, N, f->d(F_ZS), f->d(F_V1), f->d(F_S), f->d(F_ALPHA));
    gemv_t_f64(s, f->d(F_U), N, N, Mz, f->d(F_ALPHA), f->d(F_Q1));
    trsv_lower(c, s, f->d(F_LZ), Mz, f->d(F_DINVZ), Mz, f->d(F_Q1), f->d(F_Q2), 0, flags);
    trsv_lower(c, s, f->d(F_LZ), Mz, f->d(F_DINVZ), Mz, f->d(F_Q2), f->d(F_Q1), 1, flags);
    gemv_t_f64(s, f->d(F_UT), Mz, Mz, N, f->d(F_Q1), f->d(F_V2));
    vec_mul(s, n, f->d(F_D), f->d(F_ALPHA), f->d(F_V3));
    fill_f64(s, f->d(F_FMEAN), N, 0.0);
    vec_axpby(s, n, 1.0, f->d(F_V2), 1.0, f->d(F_V3), f->d(F_FMEAN));
}

// q_i = |Lc^-1 u_i|^2 for the current G (into `out`), through Lc^-1 (F_XG) and F_W
static void fitc_q(gpl_fitc* f, double* out) {
    gpl_ctx* c = f->ctx;
    hipStream_t s = c->stream;
    const int64_t N = f->N, Mz = f->Mz;
    trtri_blocked(c, f->d(F_G), f->d(F_DINVG), Mz, f->d(F_XG));
    Prof p(c, "fitc_q", (double)Mz * Mz * N, 16.0 * N * Mz);
    gemm2_f64(s, 0, 0, Mz, N, Mz, 1.0, f->d(F_XG), Mz, f->d(F_UT), Mz, 0.0, f->d(F_W), Mz, 3, 0);
    colsumsq_f64(s, f->d(F_W), Mz, Mz, f->n, out, 0);
}

// Damped parallel EP over the FITC prior: ep.method_fitc (R/ep.R:332-370; its
// tolerance is 1e-3) around the log-concave branch of ep_iter.method_fitc
// (R/ep.R:148-185,271-278).
static int fitc_ep_loop(gpl_fitc* f, const double* tr, const double* off) {
    gpl_ctx* c = f->ctx;
    hipStream_t s = c->stream;
    const int64_t n = f->n, N = f->N, Mz = f->Mz;
    const int q = f->opts.quad_order > 0 ? f->opts.quad_order : 11;
    if (q > 32) return fail(GPL_ERR_ARG, "quad_order must be <= 32");
    std::vector<double> gx, gw;
    gauss_hermite(q, gx, gw);
    HIP_TRY(hipMemcpyAsync(f->d(F_GH), gx.data(), q * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(f->d(F_GH) + 32, gw.data(), q * 8, hipMemcpyHostToDevice, s));
    // P_old = 0, z_old = 0, fmean_old = 0, fvar_old = diag K + jitter (R/gp_fit.R:171, R/ep.R:339)
    fill_f64(s, f->d(F_POLD), N, 0.0);
    fill_f64(s, f->d(F_ZOLD), N, 0.0);
    fill_f64(s, f->d(F_FOLD), N, 0.0);
    fill_f64(s, f->d(F_V3), N, 0.0);
    launch_build_diag(s, f->b[F_PROG].as<DevProg>(), f->prog, f->d(F_F), N, f->d(F_F), N, n, f->d(F_V2));
    fitc_diag(s, n, N, f->d(F_V2), f->d(F_V3), f->opts.jitter, f->d(F_FVOLD));
    const int maxiter = f->opts.maxiter > 0 ? f->opts.maxiter : 100;
    const double tol = 1e-3, damping = f->opts.damping;
    const double dmin = f->opts.damping_min > 0 ? f->opts.damping_min : 0.1;
    double damp = damping, diff_old = INFINITY, diff_new = INFINITY;
    bool quad_ok = true;
    int* flags = f->b[F_INFO].as<int>();
    int fl[2];
    double sc[8];
    int it;
    for (it = 1; it <= maxiter; ++it) {
        HIP_TRY(hipMemsetAsync(flags, 0, 8, s));
        ep_sites(s, f->lik.kind, f->lik.link, lik_par(f->lik), n, f->d(F_FOLD), f->d(F_FVOLD), f->d(F_ZOLD), f->d(F_POLD), f->d(F_Y),
                 tr, off, q, f->d(F_GH), f->d(F_GH) + 32, damp, f->d(F_Z), f->d(F_V), f->d(F_SV), f->d(F_LOGZS),
                 f->d(F_CAVM), f->d(F_CAVV), flags + 1);
        HIP_TRY(hipMemcpyAsync(fl, flags, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (fl[1] & 2) {
            f->logZ = -INFINITY;
            f->iters = it;
            return fail(GPL_ERR_NAN, "NaN in EP site parameters (R/ep.R:163-165 returns log_evidence = -Inf)");
        }
        if (fl[1] & 4) {
            f->logZ = -INFINITY;
            return fail(GPL_ERR_UNSUPPORTED, "EP site variance <= 0: R/ep.R:186-270 is not on the device path");
        }
        if (fl[1] & 1) quad_ok = false;
        fitc_lemma_step(f, flags);
        fitc_q(f, f->d(F_V1));
        fitc_ep_var(s, n, f->d(F_V), f->d(F_S), f->d(F_D), f->d(F_V1), f->d(F_FVAR));
        // per-site evidence terms (the binv argument only feeds an unused output here)
        ep_post(s, n, f->d(F_V), f->d(F_Z), f->d(F_V1), f->d(F_CAVM), f->d(F_CAVV), f->d(F_LOGZS), f->d(F_V4),
                f->d(F_TERM));
        double* scd = f->d(F_SCAL);
        reduce_sum(s, f->d(F_TERM), n, scd + 0);
        reduce_dot(s, f->d(F_Z), f->d(F_ALPHA), n, scd + 1);
        vec_log(s, n, f->d(F_V), f->d(F_V3));
        reduce_sum(s, f->d(F_V3), n, scd + 2);
        vec_log(s, n, f->d(F_S), f->d(F_V4));
        reduce_sum(s, f->d(F_V4), n, scd + 3);
        reduce_sum(s, f->d(F_LDZ), Mz, scd + 4);
        reduce_sum(s, f->d(F_LDG), Mz, scd + 5);
        reduce_maxabsdiff(s, f->d(F_FOLD), f->d(F_FMEAN), n, scd + 6);
        int r = fitc_scalars(f, sc, 7, fl);
        if (r) return r;
        if (fl[0] == kPanelWaitTimeout) return wait_timeout_fail();
        if (fl[0] > 0) {
            set_minor(fl[0]);
            return fail(GPL_ERR_NOT_PD, "inversion lemma: the leading minor of order " + std::to_string(fl[0]) +
                                            " of Kz + K_zx S^-1 K_xz is not positive");
        }
        // log Z_EP (R/ep.R:179-181); the site terms carry -1/2 log V, added back here
        const double logdet = sc[3] - 2.0 * sc[4] + 2.0 * sc[5];
        f->logZ = sc[0] + 0.5 * sc[2] - 0.5 * logdet - 0.5 * sc[1];
        diff_new = sc[6];
        if (diff_new > diff_old) {
            damp = damp * damping;
            continue;
        }
        damp = std::max(damp, dmin);
        HIP_TRY(hipMemcpyAsync(f->d(F_FOLD), f->d(F_FMEAN), N * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(f->d(F_FVOLD), f->d(F_FVAR), N * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(f->d(F_ZOLD), f->d(F_Z), N * 8, hipMemcpyDeviceToDevice, s));
        vec_recip(s, n, f->d(F_V), f->d(F_POLD));
        diff_old = diff_new;
        if (diff_new < tol) break;
    }
    if (it > maxiter) it = maxiter;
    f->iters = it;
    if (!quad_ok) f->logZ = -INFINITY;
    if (maxiter > 1 && it == maxiter) {  // R/ep.R:362-365
        f->logZ = -INFINITY;
        fail(GPL_ERR_NOCONV, "Maximum number of iterations in EP reached, max delta f = " + std::to_string(diff_new));
        return GPL_ERR_NOCONV;
    }
    return GPL_OK;
}

static int fitc_run(gpl_fitc* f) {
    gpl_ctx* c = f->ctx;
    hipStream_t s = c->stream;
    const int64_t n = f->n, N = f->N, m = f->m, Mz = f->Mz;
    const DevProg* pd = f->b[F_PROG].as<DevProg>();
    const double jitter = f->opts.jitter;
    int* flags = f->b[F_INFO].as<int>();
    int fl[2] = {0, 0};
    // Kz = K(z, z) + jitter I (padding: identity), its factor Lz and Lz^-1
    fill_f64(s, f->d(F_KZ), Mz * Mz, 0.0);
    launch_build_sym_f64(s, pd, f->prog, f->d(F_FZ), m, Mz, jitter, 0, 1.0, nullptr, f->d(F_KZ), Mz);
    HIP_TRY(hipMemcpyAsync(f->d(F_LZ), f->d(F_KZ), Mz * Mz * 8, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemsetAsync(flags, 0, 8, s));
    potrf_blocked(c, f->d(F_LZ), Mz, f->d(F_DINVZ), f->d(F_LDZ), flags);
    double sc[8];
    int r = fitc_scalars(f, sc, 1, fl);
    if (r) return r;
    if (fl[0] == kPanelWaitTimeout) return wait_timeout_fail();
    if (fl[0] > 0) {
        set_minor(fl[0]);
        return fail(GPL_ERR_NOT_PD, "Kz: the leading minor of order " + std::to_string(fl[0]) + " is not positive");
    }
    trtri_blocked(c, f->d(F_LZ), f->d(F_DINVZ), Mz, f->d(F_XZ));
    // K_xz in both orientations (zero outside n x m)
    {
        Prof p(c, "fitc_cross", 0, 16.0 * N * Mz);
        launch_build_cross(s, pd, f->prog, f->d(F_F), N, n, N, f->d(F_FZ), Mz, m, Mz, f->d(F_U), N);
        transpose_f64(s, f->d(F_U), N, N, Mz, f->d(F_UT), Mz);
    }
    // D = diag K + jitter - colSums((Lz^-1 K_zx)^2)  (R/gp_fit.R:107-108); F_W holds Lz^-1 K_zx here
    {
        Prof p(c, "fitc_diag_trmm", (double)Mz * Mz * N, 16.0 * N * Mz);
        gemm2_f64(s, 0, 0, Mz, N, Mz, 1.0, f->d(F_XZ), Mz, f->d(F_UT), Mz, 0.0, f->d(F_W), Mz, 3, 0);
        colsumsq_f64(s, f->d(F_W), Mz, Mz, n, f->d(F_V1), 0);
    }
    launch_build_diag(s, pd, f->prog, f->d(F_F), N, f->d(F_F), N, n, f->d(F_V2));
    fitc_diag(s, n, N, f->d(F_V2), f->d(F_V1), jitter, f->d(F_D));

    const double* tr = f->has_trials ? f->d(F_TR) : nullptr;
    const double* off = f->has_offset ? f->d(F_OFF) : nullptr;
    if (f->opts.approx == GPL_APPROX_EP) {
        r = fitc_ep_loop(f, tr, off);
        if (r && r != GPL_ERR_NOCONV) return r;
        gemv_t_f64(s, f->d(F_U), N, N, Mz, f->d(F_ALPHA), f->d(F_Q1));
        trsv_lower_f64(s, f->d(F_LZ), Mz, f->d(F_DINVZ), Mz, f->d(F_Q1), f->d(F_WZ), 0);
        {
            Prof p(c, "fitc_pred_matrix", 0, 0);
            fitc_predictive_matrix(f);
        }
        HIP_TRY(hipStreamSynchronize(s));
        return r;
    }
    const bool gauss = f->lik.kind == GPL_LIK_GAUSSIAN;
    const int maxiter = gauss ? 1 : (f->opts.maxiter > 0 ? f->opts.maxiter : 30);  // R/laplace.R:303-305
    double log_post_old = -INFINITY, diff = INFINITY;
    fill_f64(s, f->d(F_FOLD), N, 0.0);
    int it;
    for (it = 1; it <= maxiter; ++it) {
        HIP_TRY(hipMemsetAsync(flags, 0, 8, s));
        if (gauss)
            lik_pseudodata_la(s, GPL_LIK_GAUSSIAN, 0, f->lik.sigma, n, nullptr, f->d(F_Y), nullptr, nullptr,
                              f->d(F_Z), f->d(F_V), f->d(F_SV), flags + 1);
        else
            lik_pseudodata_la(s, f->lik.kind, f->lik.link, lik_par(f->lik), n, f->d(F_FOLD), f->d(F_Y), tr, off, f->d(F_Z),
                              f->d(F_V), f->d(F_SV), flags + 1, (f->opts.flags & GPL_FLAG_ANALYTIC_W) ? 1 : 0);
        // inversion lemma for C = K_xz Kz^-1 K_zx + S, S = D + V (R/linear_algebra.R:3-47)
        fitc_lemma_step(f, flags);
        // log evidence (R/laplace.R:143-151): logdet C - log|V| from the lemma
        lik_loglik(s, f->lik.kind, f->lik.link, lik_par(f->lik), n, f->d(F_FMEAN), f->d(F_Y), tr, off,
                   f->d(F_V3));
        double* scd = f->d(F_SCAL);
        reduce_sum(s, f->d(F_V3), n, scd + 0);
        reduce_dot(s, f->d(F_FMEAN), f->d(F_ALPHA), n, scd + 1);
        vec_log(s, n, f->d(F_V), f->d(F_V3));
        reduce_sum(s, f->d(F_V3), n, scd + 2);
        vec_log(s, n, f->d(F_S), f->d(F_V4));
        reduce_sum(s, f->d(F_V4), n, scd + 3);
        reduce_sum(s, f->d(F_LDZ), Mz, scd + 4);
        reduce_sum(s, f->d(F_LDG), Mz, scd + 5);
        r = fitc_scalars(f, sc, 6, fl);
        if (r) return r;
        if (fl[1] & 6) {
            f->logZ = -INFINITY;
            return fail((fl[1] & 2) ? GPL_ERR_NAN : GPL_ERR_UNSUPPORTED,
                        (fl[1] & 2) ? "NaN in Laplace pseudo-data"
                                    : "non-log-concave pseudo-variance (V <= 0): R/laplace.R:155-235 is not on "
                                      "the device path");
        }
        if (fl[0] == kPanelWaitTimeout) return wait_timeout_fail();
        if (fl[0] > 0) {
            set_minor(fl[0]);
            return fail(GPL_ERR_NOT_PD, "inversion lemma: the leading minor of order " + std::to_string(fl[0]) +
                                            " of Kz + K_zx S^-1 K_xz is not positive");
        }
        const double logdet = sc[3] - 2.0 * sc[4] + 2.0 * sc[5];
        const double log_post = sc[0] - 0.5 * sc[1];
        f->logZ = log_post + 0.5 * (sc[2] - logdet);
        HIP_TRY(hipMemcpyAsync(f->d(F_FOLD), f->d(F_FMEAN), N * 8, hipMemcpyDeviceToDevice, s));
        diff = log_post - log_post_old;
        log_post_old = log_post;
        if (std::fabs(diff) < f->opts.tol) break;
    }
    if (it > maxiter) it = maxiter;
    f->iters = it;
    // prediction state: w = Lz^-1 K_zx alpha and M (R/gp_pred.R:249-252,279)
    gemv_t_f64(s, f->d(F_U), N, N, Mz, f->d(F_ALPHA), f->d(F_Q1));
    trsv_lower_f64(s, f->d(F_LZ), Mz, f->d(F_DINVZ), Mz, f->d(F_Q1), f->d(F_WZ), 0);
    {
        Prof p(c, "fitc_pred_matrix", 0, 0);
        fitc_predictive_matrix(f);
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (maxiter > 1 && it == maxiter) {  // R/laplace.R:318-320 warns whenever the cap is reached
        fail(GPL_ERR_NOCONV, "Maximum number of iterations in Laplace reached, delta log post = " +
                                 std::to_string(diff));
        return GPL_ERR_NOCONV;
    }
    return GPL_OK;
}

extern "C" int gpl_fitc_fit(gpl_ctx* c, const gpl_cf_term* terms, int32_t nterms, const double* x, int64_t n,
                            int64_t ncols, const double* z, int64_t m, const double* y, const double* trials,
                            const double* offset, const gpl_lik* lik, const gpl_opts* opts, gpl_fitc** out,
                            double* log_evidence) {
    if (!c || !x || !z || !y || !lik || !out) return fail(GPL_ERR_ARG, "NULL argument");
    if (n < 1 || m < 1 || ncols < 0) return fail(GPL_ERR_ARG, "n and m must be >= 1");
    gpl_opts o;
    if (opts) o = *opts;
    else gpl_opts_default(&o);
    if (lik->kind == GPL_LIK_GAUSSIAN && o.approx == GPL_APPROX_EP)
        return fail(GPL_ERR_ARG, "approx_ep disabled for lik_gaussian");  // R/gp.R:54-56
    if (o.precision != GPL_PREC_FP64) return fail(GPL_ERR_UNSUPPORTED, "FITC runs in FP64");
    if (lik->kind == GPL_LIK_GAUSSIAN && !(lik->sigma > 0)) return fail(GPL_ERR_ARG, "sigma must be positive");
    if ((o.flags & GPL_FLAG_ANALYTIC_W) && lik->kind == GPL_LIK_BETABINOM)
        return fail(GPL_ERR_UNSUPPORTED, "the exact curvature is provided for the logit, probit and log links "
                                         "(not the beta-binomial likelihood)");
    DevProg P;
    int r = prepare_prog(terms, nterms, ncols, P);
    if (r) return r;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    gpl_fitc* f = new gpl_fitc();
    f->ctx = c;
    f->n = n;
    f->N = pad64(n);
    f->m = m;
    f->Mz = pad64(m);
    f->ncols = ncols;
    f->prog = P;
    f->lik = *lik;
    f->opts = o;
    f->has_trials = trials != nullptr;
    f->has_offset = offset != nullptr;
    r = fitc_alloc(f);
    hipStream_t s = c->stream;
    auto setup = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(f->b[F_PROG].p, &f->prog, sizeof(DevProg), hipMemcpyHostToDevice, s));
        if (n * ncols > 0) HIP_TRY(hipMemcpyAsync(f->d(F_X), x, n * ncols * 8, hipMemcpyHostToDevice, s));
        if (m * ncols > 0) HIP_TRY(hipMemcpyAsync(f->d(F_ZX), z, m * ncols * 8, hipMemcpyHostToDevice, s));
        const DevProg* pd = f->b[F_PROG].as<DevProg>();
        launch_features(s, f->d(F_X), n, n, (int)ncols, pd, f->prog, f->d(F_F), f->N);
        launch_features(s, f->d(F_ZX), m, m, (int)ncols, pd, f->prog, f->d(F_FZ), f->Mz);
        int q = fitc_upload(f, F_Y, y);
        if (!q) q = fitc_upload(f, F_TR, trials);
        if (!q) q = fitc_upload(f, F_OFF, offset);
        return q;
    };
    if (!r) r = setup();
    if (!r) r = fitc_run(f);
    if (r && r != GPL_ERR_NOCONV) {
        const double lz = f->logZ;
        delete f;
        if (log_evidence) *log_evidence = (r == GPL_ERR_NAN) ? -INFINITY : lz;
        return r;
    }
    *out = f;
    if (log_evidence) *log_evidence = f->logZ;
    return r;
}

extern "C" int gpl_fitc_pred(gpl_fitc* f, const double* xt, int64_t mt, int32_t want_var, double* mean,
                             double* var) {
    if (!f || (!xt && mt > 0) || !mean || (want_var && !var)) return fail(GPL_ERR_ARG, "NULL argument");
    if (mt <= 0) return GPL_OK;
    gpl_ctx* c = f->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t m = f->m, Mz = f->Mz, Mt = pad64(mt);
    const DevProg* pd = f->b[F_PROG].as<DevProg>();
    double* xtd = (double*)c->buf(10, (size_t)mt * std::max<int64_t>(f->ncols, 1) * 8);
    double* Ft = (double*)c->buf(11, (size_t)Mt * MAX_FEAT * 8);
    double* KzT = (double*)c->buf(12, (size_t)Mz * Mt * 8);
    double* A2 = (double*)c->buf(13, (size_t)Mz * Mt * 8);
    double* Yv = (double*)c->buf(14, (size_t)Mz * Mt * 8);
    double* vb = (double*)c->buf(15, (size_t)Mt * 6 * 8);
    if (!xtd || !Ft || !KzT || !A2 || !Yv || !vb) return fail(GPL_ERR_OOM, "device allocation failed");
    if (mt * f->ncols > 0) HIP_TRY(hipMemcpyAsync(xtd, xt, mt * f->ncols * 8, hipMemcpyHostToDevice, s));
    launch_features(s, xtd, mt, mt, (int)f->ncols, pd, f->prog, Ft, Mt);
    launch_build_cross(s, pd, f->prog, f->d(F_FZ), Mz, m, Mz, Ft, Mt, mt, Mt, KzT, Mz);
    // A_t = Lz^-1 K_zt (Kt_aux2); mean = A_t^T w
    gemm2_f64(s, 0, 0, Mz, Mt, Mz, 1.0, f->d(F_XZ), Mz, KzT, Mz, 0.0, A2, Mz, 3, 0);
    gemv_t_f64(s, A2, Mz, Mz, mt, f->d(F_WZ), vb);
    if (want_var) {
        // prior_var = colSums(K_zt * Kz^-1 K_zt) + Dt, Dt = diag K_tt - colSums(A_t^2)  (R/gp_pred.R:244-246,271)
        double* c1 = vb + Mt;
        double* c2 = vb + 2 * Mt;
        double* kd = vb + 3 * Mt;
        double* red = vb + 4 * Mt;
        gemm2_f64(s, 1, 0, Mz, Mt, Mz, 1.0, f->d(F_XZ), Mz, A2, Mz, 0.0, Yv, Mz, 0, 0);  // Kz^-1 K_zt = Lz^-T A_t
        coldot_f64(s, KzT, Yv, Mz, Mz, mt, c1);
        colsumsq_f64(s, A2, Mz, Mz, mt, c2, 0);
        launch_build_diag(s, pd, f->prog, Ft, Mt, Ft, Mt, mt, kd);
        vec_axpby(s, mt, 1.0, kd, -1.0, c2, c2);   // Dt
        vec_axpby(s, mt, 1.0, c1, 1.0, c2, c2);    // prior_var
        // var_reduction = colSums(A_t * (M A_t))  (R/gp_pred.R:278-280)
        gemm2_f64(s, 0, 0, Mz, Mt, Mz, 1.0, f->d(F_MM), Mz, A2, Mz, 0.0, Yv, Mz, 0, 0);
        coldot_f64(s, A2, Yv, Mz, Mz, mt, red);
        vec_axpby(s, mt, 1.0, c2, -1.0, red, red);
        HIP_TRY(hipMemcpyAsync(var, red, mt * 8, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipMemcpyAsync(mean, vb, mt * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GPL_OK;
}

// gp_draw_analytic.method_fitc with marginal = FALSE (R/gp_draw.R:126-140):
// mean + chol(cov + jitter I) z, cov = A_t^T A_t + diag(Dt) - A_t^T M A_t
// (R/gp_pred.R:258-268, train = FALSE).
extern "C" int gpl_fitc_draw(gpl_fitc* f, const double* xt, int64_t mt, double jitter, const double* z,
                             int64_t draws, double* out) {
    if (!f || !xt || !z || !out || mt < 1 || draws < 1) return fail(GPL_ERR_ARG, "bad argument");
    gpl_ctx* c = f->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t m = f->m, Mz = f->Mz, Mt = pad64(mt), Dp = pad64(draws);
    const DevProg* pd = f->b[F_PROG].as<DevProg>();
    double* xtd = (double*)c->buf(10, (size_t)mt * std::max<int64_t>(f->ncols, 1) * 8);
    double* Ft = (double*)c->buf(11, (size_t)Mt * MAX_FEAT * 8);
    double* KzT = (double*)c->buf(12, (size_t)Mz * Mt * 8);
    double* A2 = (double*)c->buf(13, (size_t)Mz * Mt * 8);
    double* Yv = (double*)c->buf(14, (size_t)Mz * Mt * 8);
    double* vb = (double*)c->buf(15, (size_t)Mt * 6 * 8);
    double* Cv = (double*)c->buf(16, (size_t)Mt * Mt * 8);
    double* Dv = (double*)c->buf(17, (size_t)Mt * 64 * 8 + (size_t)Mt * 8 + 64);
    double* Zd = (double*)c->buf(18, (size_t)Mt * Dp * 8);
    double* Wd = (double*)c->buf(19, (size_t)Mt * Dp * 8);
    if (!xtd || !Ft || !KzT || !A2 || !Yv || !vb || !Cv || !Dv || !Zd || !Wd)
        return fail(GPL_ERR_OOM, "device allocation failed");
    double* ld = Dv + Mt * 64;
    int* info = (int*)(ld + Mt);
    if (mt * f->ncols > 0) HIP_TRY(hipMemcpyAsync(xtd, xt, mt * f->ncols * 8, hipMemcpyHostToDevice, s));
    launch_features(s, xtd, mt, mt, (int)f->ncols, pd, f->prog, Ft, Mt);
    launch_build_cross(s, pd, f->prog, f->d(F_FZ), Mz, m, Mz, Ft, Mt, mt, Mt, KzT, Mz);
    gemm2_f64(s, 0, 0, Mz, Mt, Mz, 1.0, f->d(F_XZ), Mz, KzT, Mz, 0.0, A2, Mz, 3, 0);
    gemv_t_f64(s, A2, Mz, Mz, mt, f->d(F_WZ), vb);
    // Dt = diag K_tt - colSums(A_t^2)
    double* dt = vb + Mt;
    colsumsq_f64(s, A2, Mz, Mz, mt, dt, 0);
    launch_build_diag(s, pd, f->prog, Ft, Mt, Ft, Mt, mt, vb + 2 * Mt);
    vec_axpby(s, mt, 1.0, vb + 2 * Mt, -1.0, dt, dt);
    // lower tiles of A_t^T A_t - A_t^T (M A_t), then the diagonal terms
    gemm2_f64(s, 0, 0, Mz, Mt, Mz, 1.0, f->d(F_MM), Mz, A2, Mz, 0.0, Yv, Mz, 0, 0);
    gemm2_f64(s, 1, 0, Mt, Mt, Mz, 1.0, A2, Mz, A2, Mz, 0.0, Cv, Mt, 1, 0);
    gemm2_f64(s, 1, 0, Mt, Mt, Mz, -1.0, A2, Mz, Yv, Mz, 1.0, Cv, Mt, 1, 0);
    add_diag_f64(s, Cv, Mt, mt, Mt, dt, jitter);
    HIP_TRY(hipMemsetAsync(info, 0, 64, s));
    potrf_blocked(c, Cv, Mt, Dv, ld, info);
    int inf = 0;
    HIP_TRY(hipMemcpyAsync(&inf, info, 4, hipMemcpyDeviceToHost, s));
    fill_f64(s, Zd, Mt * Dp, 0.0);
    HIP_TRY(upload_2d(c, s, Zd, Mt, z, mt, draws, 20));
    gemm2_f64(s, 0, 0, Mt, Dp, Mt, 1.0, Cv, Mt, Zd, Mt, 0.0, Wd, Mt, 3, 0);
    std::vector<double> mean(mt);
    HIP_TRY(hipMemcpyAsync(mean.data(), vb, mt * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(download_2d(c, s, out, Wd, Mt, mt, draws, 21));
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

// Predictive variance at the training inputs with the FITC diagonal
// (gp_pred_post.method_fitc, train = TRUE: R/gp_pred.R:270-285), the input of
// loo_posteriors.approx_laplace (R/gp_loo.R:120-130).  With A = Lz^-1 K_zx,
// Dt = diag K - colSums(A^2) and q_i = |Lc^-1 u_i|^2 the reference's three terms
// are (A^T M A)_ii, 2 Dt_i q_i / S_i and Dt_i^2 [C^-1]_ii.
extern "C" int gpl_fitc_train_var(gpl_fitc* f, double* out) {
    if (!f || !out) return fail(GPL_ERR_ARG, "NULL argument");
    gpl_ctx* c = f->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = f->n, N = f->N, Mz = f->Mz;
    double* A = (double*)c->buf(22, (size_t)Mz * N * 8);
    double* B = (double*)c->buf(23, (size_t)Mz * N * 8);
    double* v = (double*)c->buf(24, (size_t)N * 6 * 8);
    double* X = (double*)c->buf(25, (size_t)Mz * Mz * 8);
    if (!A || !B || !v || !X) return fail(GPL_ERR_OOM, "device allocation failed");
    double *c2 = v, *kd = v + N, *c1 = v + 2 * N, *t1 = v + 3 * N, *q = v + 4 * N, *res = v + 5 * N;
    gemm2_f64(s, 0, 0, Mz, N, Mz, 1.0, f->d(F_XZ), Mz, f->d(F_UT), Mz, 0.0, A, Mz, 3, 0);  // A = Lz^-1 K_zx
    colsumsq_f64(s, A, Mz, Mz, n, c2, 0);
    launch_build_diag(s, f->b[F_PROG].as<DevProg>(), f->prog, f->d(F_F), N, f->d(F_F), N, n, kd);
    vec_axpby(s, n, 1.0, kd, -1.0, c2, c2);  // Dt
    gemm2_f64(s, 1, 0, Mz, N, Mz, 1.0, f->d(F_XZ), Mz, A, Mz, 0.0, B, Mz, 0, 0);  // Kz^-1 K_zx
    coldot_f64(s, f->d(F_UT), B, Mz, Mz, n, c1);
    vec_axpby(s, n, 1.0, c1, 1.0, c2, c1);  // prior_var
    gemm2_f64(s, 0, 0, Mz, N, Mz, 1.0, f->d(F_MM), Mz, A, Mz, 0.0, B, Mz, 0, 0);  // M A
    coldot_f64(s, A, B, Mz, Mz, n, t1);
    trtri_blocked(c, f->d(F_G), f->d(F_DINVG), Mz, X);
    gemm2_f64(s, 0, 0, Mz, N, Mz, 1.0, X, Mz, f->d(F_UT), Mz, 0.0, B, Mz, 3, 0);  // Lc^-1 K_zx
    colsumsq_f64(s, B, Mz, Mz, n, q, 0);
    fitc_train_var(s, n, c1, t1, c2, q, f->d(F_S), res);
    HIP_TRY(hipMemcpyAsync(out, res, n * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GPL_OK;
}

extern "C" int gpl_fitc_vector(gpl_fitc* f, int32_t which, double* out) {
    if (!f || !out) return fail(GPL_ERR_ARG, "NULL argument");
    int idx;
    switch (which) {
        case GPL_VEC_FMEAN: idx = F_FMEAN; break;
        case GPL_VEC_ALPHA: idx = F_ALPHA; break;
        case GPL_VEC_Z: idx = F_Z; break;
        case GPL_VEC_V: idx = F_V; break;
        case GPL_VEC_FVAR: idx = F_FVAR; break;
        case GPL_VEC_CAVITY_MEAN: idx = F_CAVM; break;
        case GPL_VEC_CAVITY_VAR: idx = F_CAVV; break;
        case 7: idx = F_D; break;
        default: return fail(GPL_ERR_ARG, "unknown FITC vector");
    }
    std::lock_guard<std::mutex> lk(f->ctx->mu);
    HIP_TRY(hipSetDevice(f->ctx->device));
    HIP_TRY(hipMemcpyAsync(out, f->d(idx), f->n * 8, hipMemcpyDeviceToHost, f->ctx->stream));
    HIP_TRY(hipStreamSynchronize(f->ctx->stream));
    return GPL_OK;
}

extern "C" int gpl_fitc_info(gpl_fitc* f, int64_t* n, int64_t* m, int32_t* iters, double* log_evidence) {
    if (!f) return fail(GPL_ERR_ARG, "NULL model");
    if (n) *n = f->n;
    if (m) *m = f->m;
    if (iters) *iters = f->iters;
    if (log_evidence) *log_evidence = f->logZ;
    return GPL_OK;
}

extern "C" void gpl_fitc_free(gpl_fitc* f) {
    if (!f) return;
    hipSetDevice(f->ctx->device);
    hipStreamSynchronize(f->ctx->stream);
    delete f;
}
