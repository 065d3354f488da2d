// The device model behind a gpl_model handle (shared by the C-ABI translation units).
#pragma once

#include "host_common.h"

extern "C" void gpl_model_free(gpl_model* m);

enum {
    B_X, B_F, B_Y, B_TR, B_OFF, B_L, B_DINV, B_LDIAG, B_K, B_XINV, B_BINV, B_PROG,
    B_Z, B_V, B_SV, B_ALPHA, B_FMEAN, B_FVAR, B_CAVM, B_CAVV, B_LOGZS, B_TERM,
    B_FOLD, B_FVOLD, B_ZOLD, B_POLD, B_W1, B_W2, B_W3, B_W4, B_SCAL, B_PART, B_KMV, B_GH, B_INFO, B_DACC, B_B64,
    // non-log-concave branch (split sites): site order, S1 = V_good^-1/2, L1^-1 S1 K_gb,
    // chol of the bad sites' predictive covariance (Lk) and of A = I + Lk^T V_bad^-1 Lk,
    // their diagonal-block inverses and log pivots, 1/V_bad, vector and matrix workspace
    B_SIDX, B_S1, B_AUX, B_LK, B_DK, B_LDK, B_AC, B_DA, B_LDA, B_IV2, B_SPV, B_SPG, B_SPH, B_SPQ,
    B_COUNT
};

struct gpl_model {
    gpl_ctx* ctx = nullptr;
    int64_t n = 0, N = 0, ncols = 0;
    gpl::DevProg prog{};
    gpl_lik lik{};
    gpl_opts opts{};
    bool has_trials = false, has_offset = false;
    bool have_xinv = false, have_binv = false;
    double round_corr = 0.0;  // mixed form: first-order log|B| correction applied to logZ
    double logZ = 0;
    int32_t iters = 0;
    int32_t refine_iters_done = 0;
    double refine_rel = 0;
    bool refit_fp64 = false;  // mixed form redone in FP64 (refinement did not contract / FP32 pivot failed)
    // The last Laplace step / EP sweep split its sites (some pseudo-variance <= 0):
    // the factor in B_L (leading dimension Ng) is that of the ng good sites, and the
    // posterior of the nb bad ones is held by B_AUX, B_LK, B_AC (the reference's
    // C_lu, R/laplace.R:40-117, R/ep.R:50-138).
    bool split = false;
    int64_t ng = 0, nb = 0, Ng = 0, Nb = 0;
    bool no_posterior = false;  // the split's A was not positive definite: log evidence -Inf
    bool restored = false;      // built by gpl_model_restore: the fit fields only, no targets
    DevBuf b[B_COUNT];
    ~gpl_model() {
        for (auto& x : b) x.free_();
    }
    double* d(int i) const { return b[i].as<double>(); }
};

#define ENSURE(m, idx, bytes)                                                             \
    do {                                                                                  \
        if (!(m)->b[idx].ensure(bytes)) return fail(GPL_ERR_OOM, "device allocation failed"); \
    } while (0)


namespace gpl {
// Non-log-concave branch (split.cpp): one Laplace step (ep = false) or EP sweep
// on split sites; and the predictive variance reduction of a split fit.
int split_step(gpl_model* m, bool ep, double* logZ);
int split_var_reduction(gpl_ctx* c, gpl_model* m, const double* KtT, int64_t M, double* red);
// the buffers of a Gaussian-likelihood model of n points (factor FP32 when mixed)
int model_alloc(gpl_model* m, int64_t n, int64_t ncols, bool mixed);
gpl_model* model_new(gpl_ctx* c);
}  // namespace gpl
