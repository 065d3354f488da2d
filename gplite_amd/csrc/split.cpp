// Non-log-concave branch of the Laplace step and the EP sweep on the device.
//
// When some pseudo-variances V_i are not positive (a beta-binomial outlier
// makes the log density locally convex), B = I + S K S has no real S and the
// reference splits the sites (R/laplace.R:40-117, R/ep.R:50-138): the "good"
// sites (V > 0) are fitted as usual, the "bad" ones get the Gaussian predictive
// N(m, P) of their latent values given the good pseudo-data, the evidence is
// p(y_good) p(y_bad | y_good), and the posterior solves (K + V) alpha = z with an
// LU factorisation of the indefinite K + V.
//
// Here the LU is replaced by block elimination over the same split, with every
// piece a positive definite factor the dense kernels already provide:
//   (K + V) = [[M11, K12], [K21, K22 + V2]],  M11 = K11 + V1 = C1 C1^T,
//   Schur complement S = K22 + V2 - K21 M11^-1 K12 = P + V2,  P = Lk Lk^T,
//   S^-1 = V2^-1 - V2^-1 Lk A^-1 Lk^T V2^-1,  A = I + Lk^T V2^-1 Lk
// (Woodbury), where Lk and chol(A) are the factors the reference forms for the
// evidence.  alpha, K alpha, the EP posterior variances diag(V (K + V)^-1 K)
// and the predictive variances k^T (K + V)^-1 k follow from these in exact
// arithmetic; they agree with the LU solution to rounding.
#include <algorithm>
#include <cmath>
#include <vector>

#include "host_common.h"
#include "model.h"

using namespace gpl;

namespace {

// B = L^-1 B in place for the N x M panel B (leading dimension ldb): the blocked
// forward substitution of the prediction path, with the diagonal-block inverses.
void trsm_lower(hipStream_t s, const double* L, int64_t ld, const double* Dinv, int64_t N, double* B, int64_t ldb,
                int64_t M) {
    const int64_t T = N / 64;
    for (int64_t i = 0; i < T; ++i) {
        gemm2_f64(s, 0, 0, 64, M, 64, 1.0, Dinv + i * 64 * 64, 64, B + i * 64, ldb, 0.0, B + i * 64, ldb, 0, 0);
        if (i + 1 < T)
            gemm2_f64(s, 0, 0, (T - 1 - i) * 64, M, 64, -1.0, L + (i + 1) * 64 + i * 64 * ld, ld, B + i * 64, ldb, 1.0,
                      B + (i + 1) * 64, ldb, 0, 0);
    }
}

// Workspace vectors carved from B_SPV (each Kmax long, Kmax = max(Ng, Nb, N)).
struct Vecs {
    double* v[16];
};

int ensure_split(gpl_model* m) {
    const int64_t Ng = m->Ng, Nb = m->Nb, N = m->N;
    const int64_t Km = std::max(std::max(Ng, Nb), N);
    ENSURE(m, B_SIDX, (size_t)m->n * 8);
    ENSURE(m, B_S1, (size_t)Ng * 8);
    ENSURE(m, B_AUX, (size_t)Ng * Nb * 8);
    ENSURE(m, B_LK, (size_t)Nb * Nb * 8);
    ENSURE(m, B_DK, (size_t)Nb * 64 * 8);
    ENSURE(m, B_LDK, (size_t)Nb * 8);
    ENSURE(m, B_AC, (size_t)Nb * Nb * 8);
    ENSURE(m, B_DA, (size_t)Nb * 64 * 8);
    ENSURE(m, B_LDA, (size_t)Nb * 8);
    ENSURE(m, B_IV2, (size_t)Nb * 8);
    ENSURE(m, B_SPV, (size_t)Km * 16 * 8);
    ENSURE(m, B_SPQ, (size_t)Nb * Nb * 8);
    return GPL_OK;
}

// alpha-type solve with the good sites' factor: out = S1 L1^-T L1^-1 (S1 v)
void solve_good(gpl_model* m, hipStream_t s, const double* v, double* out, double* t1, double* t2) {
    const int64_t Ng = m->Ng;
    vec_mul(s, Ng, m->d(B_S1), v, t1);
    trsv_lower_f64(s, m->d(B_L), Ng, m->d(B_DINV), Ng, t1, t2, 0);
    trsv_lower_f64(s, m->d(B_L), Ng, m->d(B_DINV), Ng, t2, t1, 1);
    vec_mul(s, Ng, m->d(B_S1), t1, out);
}

// out = S^-1 r for the bad sites' Schur complement (all Nb long, padding zero):
// t = r / V2, out = t - V2^-1 Lk A^-1 Lk^T t
void solve_schur(gpl_model* m, hipStream_t s, const double* r, double* out, double* t, double* h, double* u) {
    const int64_t Nb = m->Nb;
    vec_mul(s, Nb, m->d(B_IV2), r, t);
    trmv_lower_f64(s, m->d(B_LK), Nb, Nb, t, h, 1, nullptr);
    trsv_lower_f64(s, m->d(B_AC), Nb, m->d(B_DA), Nb, h, u, 0);
    trsv_lower_f64(s, m->d(B_AC), Nb, m->d(B_DA), Nb, u, h, 1);
    trmv_lower_f64(s, m->d(B_LK), Nb, Nb, h, u, 0, nullptr);
    vec_mul(s, Nb, m->d(B_IV2), u, u);
    vec_axpby(s, Nb, 1.0, t, -1.0, u, out);
}

}  // namespace

// One split step.  Inputs: z, V (B_Z, B_V) at the current iterate, K in B_K.
// Laplace (ep = false): writes alpha, f = K alpha and the evidence
// (R/laplace.R:40-117).  EP: also needs the site terms (logZs, cavity moments)
// and writes fvar (R/ep.R:50-138).  Returns GPL_OK with *logZ set, GPL_ERR_NOT_PD
// when the good sites' B or the bad sites' predictive covariance is not positive
// definite (the reference's chol() raises there), or sets m->no_posterior and
// *logZ = -Inf when A is not positive definite (the reference returns
// log_evidence = -Inf there).
int gpl::split_step(gpl_model* m, bool ep, double* logZ) {
    gpl_ctx* c = m->ctx;
    hipStream_t s = c->stream;
    const int64_t n = m->n, N = m->N;
    std::vector<double> V(n);
    HIP_TRY(hipMemcpyAsync(V.data(), m->d(B_V), n * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<int64_t> idx;
    idx.reserve(n);
    for (int64_t i = 0; i < n; ++i)
        if (V[i] > 0) idx.push_back(i);
    const int64_t ng = (int64_t)idx.size(), nb = n - ng;
    for (int64_t i = 0; i < n; ++i)
        if (!(V[i] > 0)) idx.push_back(i);
    if (ng == 0) return fail(GPL_ERR_UNSUPPORTED, "every pseudo-variance is non-positive: no site to condition on");
    m->ng = ng;
    m->nb = nb;
    m->Ng = pad64(ng);
    m->Nb = pad64(nb);
    const int64_t Ng = m->Ng, Nb = m->Nb;
    int r = ensure_split(m);
    if (r) return r;
    const int64_t Km = std::max(std::max(Ng, Nb), N);
    Vecs W;
    for (int k = 0; k < 16; ++k) W.v[k] = m->d(B_SPV) + k * Km;
    HIP_TRY(hipMemcpyAsync(m->b[B_SIDX].p, idx.data(), n * 8, hipMemcpyHostToDevice, s));
    const int64_t* gi = m->b[B_SIDX].as<int64_t>();
    const int64_t* bi = gi + ng;
    int* info = m->b[B_INFO].as<int>();
    HIP_TRY(hipMemsetAsync(info + 2, 0, 8, s));
    const double* K = m->d(B_K);
    const double par = lik_par(m->lik);
    const bool tr = m->has_trials, of = m->has_offset;

    // site vectors of both groups
    double *z1 = W.v[0], *z2 = W.v[1], *V1 = W.v[2], *V2 = W.v[3], *y1 = W.v[4], *y2 = W.v[5];
    double *t1 = W.v[6], *t2 = W.v[7], *t3 = W.v[8], *a1 = W.v[9], *pm = W.v[10], *f2 = W.v[11];
    double *full = W.v[12], *full2 = W.v[13], *w1 = W.v[14], *w2 = W.v[15];
    gather_vec_f64(s, m->d(B_Z), gi, ng, Ng, z1);
    gather_vec_f64(s, m->d(B_V), gi, ng, Ng, V1);
    gather_vec_f64(s, m->d(B_Z), bi, nb, Nb, z2);
    gather_vec_f64(s, m->d(B_V), bi, nb, Nb, V2);
    rsqrt_n_f64(s, V1, ng, Ng, m->d(B_S1));
    recip_n_f64(s, V2, nb, Nb, m->d(B_IV2));

    // good sites: B1 = I + S1 K11 S1 in the factor buffer (leading dimension Ng)
    double* L1 = m->d(B_L);
    gather_sym_f64(s, K, N, gi, ng, gi, ng, m->d(B_S1), m->d(B_S1), 1, 1, L1, Ng, Ng, Ng);
    potrf_blocked(c, L1, Ng, m->d(B_DINV), m->d(B_LDIAG), info + 2);
    // alpha1 = (K11 + V1)^-1 z1; f1 and the predictive mean of the bad sites from K alpha1~
    solve_good(m, s, z1, a1, t1, t2);
    fill_f64(s, full, N, 0.0);
    scatter_vec_f64(s, a1, gi, ng, full);
    symv_lower_f64(s, K, N, N, full, full2, nullptr);
    gather_vec_f64(s, full2, gi, ng, Ng, t3);  // f1
    gather_vec_f64(s, full2, bi, nb, Nb, pm);  // pred_mean = K21 alpha1
    double* sc = m->d(B_SCAL);
    if (!ep) {
        // log p(y_good) - 1/2 f1^T alpha1 - sum log L1_ii
        double* tg = tr ? w1 : nullptr;
        double* og = of ? w2 : nullptr;
        gather_vec_f64(s, m->d(B_Y), gi, ng, Ng, y1);
        if (tg) gather_vec_f64(s, m->d(B_TR), gi, ng, Ng, tg);
        if (og) gather_vec_f64(s, m->d(B_OFF), gi, ng, Ng, og);
        lik_loglik(s, m->lik.kind, m->lik.link, par, ng, t3, y1, tg, og, t1);
        reduce_sum(s, t1, ng, sc + 10);
        reduce_dot(s, t3, a1, ng, sc + 11);
    } else {
        // EP evidence of the good sites: sum of site terms - sum log L1_ii - 1/2 z1^T alpha1
        gather_vec_f64(s, m->d(B_LOGZS), gi, ng, Ng, t1);
        gather_vec_f64(s, m->d(B_CAVM), gi, ng, Ng, w1);
        gather_vec_f64(s, m->d(B_CAVV), gi, ng, Ng, w2);
        ep_split_terms(s, ng, 0, V1, z1, w1, w2, t1, t1);
        reduce_sum(s, t1, ng, sc + 10);
        reduce_dot(s, z1, a1, ng, sc + 11);
    }
    reduce_sum(s, m->d(B_LDIAG), ng, sc + 12);

    // aux = L1^-1 S1 K12; P = K22 - aux^T aux = Lk Lk^T; A = I + Lk^T V2^-1 Lk = La La^T
    double* aux = m->d(B_AUX);
    gather_sym_f64(s, K, N, gi, ng, bi, nb, m->d(B_S1), nullptr, 0, 0, aux, Ng, Ng, Nb);
    trsm_lower(s, L1, Ng, m->d(B_DINV), Ng, aux, Ng, Nb);
    double* Lk = m->d(B_LK);
    gather_sym_f64(s, K, N, bi, nb, bi, nb, nullptr, nullptr, 0, 1, Lk, Nb, Nb, Nb);
    gemm2_f64(s, 1, 0, Nb, Nb, Ng, -1.0, aux, Ng, aux, Ng, 1.0, Lk, Nb, 1, 0);
    potrf_blocked(c, Lk, Nb, m->d(B_DK), m->d(B_LDK), info + 3);
    tril_f64(s, Lk, Nb, Nb);
    int fl[2];
    HIP_TRY(hipMemcpyAsync(fl, info + 2, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (fl[0] == kPanelWaitTimeout || fl[1] == kPanelWaitTimeout) return wait_timeout_fail();
    if (fl[0] > 0) {
        set_minor(fl[0]);
        return fail(GPL_ERR_NOT_PD, "sites with positive pseudo-variance: the leading minor of order " +
                                        std::to_string(fl[0]) + " is not positive");
    }
    if (fl[1] > 0) {
        set_minor(fl[1]);
        return fail(GPL_ERR_NOT_PD, "predictive covariance of the sites with pseudo-variance <= 0: the leading "
                                    "minor of order " + std::to_string(fl[1]) + " is not positive");
    }
    double* Tq = m->d(B_SPQ);
    double* Ac = m->d(B_AC);
    scale_rows_copy_f64(s, Lk, Nb, Nb, Nb, m->d(B_IV2), Tq);
    gemm2_f64(s, 1, 0, Nb, Nb, Nb, 1.0, Lk, Nb, Tq, Nb, 0.0, Ac, Nb, 1, 0);
    fill_f64(s, t1, Nb, 0.0);
    add_diag_f64(s, Ac, Nb, nb, Nb, t1, 1.0);
    HIP_TRY(hipMemsetAsync(info + 2, 0, 4, s));
    potrf_blocked(c, Ac, Nb, m->d(B_DA), m->d(B_LDA), info + 2);
    tril_f64(s, Ac, Nb, Nb);
    HIP_TRY(hipMemcpyAsync(fl, info + 2, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (fl[0] == kPanelWaitTimeout) return wait_timeout_fail();
    if (fl[0] > 0) {  // the posterior of all the data is not proper
        m->no_posterior = true;
        *logZ = -INFINITY;
        return GPL_OK;
    }
    m->no_posterior = false;
    reduce_sum(s, m->d(B_LDA), nb, sc + 13);  // 1/2 log|A|

    if (!ep) {
        // posterior mean of the bad sites' latent values f2 = P (alpha2 + aux_v), with
        // alpha2 = Lk^-T A^-1 Lk^T V2^-1 z2 and aux_v = Lk^-T A^-1 Lk^T P^-1 m (R/laplace.R:89-93)
        vec_mul(s, Nb, m->d(B_IV2), z2, t1);
        trmv_lower_f64(s, Lk, Nb, Nb, t1, t2, 1, nullptr);
        trsv_lower_f64(s, Ac, Nb, m->d(B_DA), Nb, t2, t1, 0);
        trsv_lower_f64(s, Ac, Nb, m->d(B_DA), Nb, t1, t2, 1);
        trsv_lower_f64(s, Lk, Nb, m->d(B_DK), Nb, t2, w1, 1);  // w1 = alpha2
        // aux_v = Lk^-T A^-1 Lk^T (P^-1 m)
        vec_axpby(s, Nb, 1.0, pm, 0.0, nullptr, t1);
        trsv_lower_f64(s, Lk, Nb, m->d(B_DK), Nb, t1, t2, 0);
        trsv_lower_f64(s, Lk, Nb, m->d(B_DK), Nb, t2, t1, 1);
        trmv_lower_f64(s, Lk, Nb, Nb, t1, t2, 1, nullptr);
        trsv_lower_f64(s, Ac, Nb, m->d(B_DA), Nb, t2, t1, 0);
        trsv_lower_f64(s, Ac, Nb, m->d(B_DA), Nb, t1, t2, 1);
        trsv_lower_f64(s, Lk, Nb, m->d(B_DK), Nb, t2, w2, 1);  // w2 = aux_v
        vec_axpby(s, Nb, 1.0, w1, 1.0, w2, t1);
        trmv_lower_f64(s, Lk, Nb, Nb, t1, t2, 1, nullptr);
        trmv_lower_f64(s, Lk, Nb, Nb, t2, f2, 0, nullptr);  // f2hat = P (alpha2 + aux_v)
        // |Lk^-1 (f2 - m)|^2
        vec_axpby(s, Nb, 1.0, f2, -1.0, pm, t1);
        trsv_lower_f64(s, Lk, Nb, m->d(B_DK), Nb, t1, t2, 0);
        reduce_dot(s, t2, t2, nb, sc + 14);
        double* tb = tr ? w1 : nullptr;
        double* ob = of ? w2 : nullptr;
        gather_vec_f64(s, m->d(B_Y), bi, nb, Nb, y2);
        if (tb) gather_vec_f64(s, m->d(B_TR), bi, nb, Nb, tb);
        if (ob) gather_vec_f64(s, m->d(B_OFF), bi, nb, Nb, ob);
        lik_loglik(s, m->lik.kind, m->lik.link, par, nb, f2, y2, tb, ob, t1);
        reduce_sum(s, t1, nb, sc + 15);
    } else {
        // (m - z2)^T Lk^-T A^-1 Lk^T V2^-1 (m - z2) (R/ep.R:99-102)
        vec_axpby(s, Nb, 1.0, pm, -1.0, z2, f2);
        vec_mul(s, Nb, m->d(B_IV2), f2, t1);
        trmv_lower_f64(s, Lk, Nb, Nb, t1, t2, 1, nullptr);
        trsv_lower_f64(s, Ac, Nb, m->d(B_DA), Nb, t2, t1, 0);
        trsv_lower_f64(s, Ac, Nb, m->d(B_DA), Nb, t1, t2, 1);
        trsv_lower_f64(s, Lk, Nb, m->d(B_DK), Nb, t2, t1, 1);
        reduce_dot(s, f2, t1, nb, sc + 14);
        // bad sites' terms: logZ + 1/2 log((c + V) / V) + 1/2 (m_c - z)^2 / (c + V)
        gather_vec_f64(s, m->d(B_LOGZS), bi, nb, Nb, t1);
        gather_vec_f64(s, m->d(B_CAVM), bi, nb, Nb, w1);
        gather_vec_f64(s, m->d(B_CAVV), bi, nb, Nb, w2);
        ep_split_terms(s, nb, 1, V2, z2, w1, w2, t1, t1);
        reduce_sum(s, t1, nb, sc + 15);
    }

    // alpha = (K + V)^-1 z by block elimination: x2 = S^-1 (z2 - m), x1 = alpha1 - M11^-1 K12 x2
    vec_axpby(s, Nb, 1.0, z2, -1.0, pm, t1);
    solve_schur(m, s, t1, f2, t2, t3, w1);  // f2 = x2
    fill_f64(s, full, N, 0.0);
    scatter_vec_f64(s, f2, bi, nb, full);
    symv_lower_f64(s, K, N, N, full, full2, nullptr);
    gather_vec_f64(s, full2, gi, ng, Ng, t1);  // K12 x2
    solve_good(m, s, t1, t2, t3, w1);           // M11^-1 K12 x2
    vec_axpby(s, Ng, 1.0, a1, -1.0, t2, t3);    // x1
    fill_f64(s, m->d(B_ALPHA), N, 0.0);
    scatter_vec_f64(s, t3, gi, ng, m->d(B_ALPHA));
    scatter_vec_f64(s, f2, bi, nb, m->d(B_ALPHA));
    symv_lower_f64(s, K, N, N, m->d(B_ALPHA), m->d(B_FMEAN), nullptr);

    if (ep) {
        // fvar_i = V_i (1 - V_i [(K + V)^-1]_ii) with
        //   good i: [M11^-1]_ii + s_i^2 g_i^T S^-1 g_i,  g_i = column i of aux^T L1^-1
        //   bad j:  [S^-1]_jj = 1/V2_j - |A^-1/2 Lk^T e_j / V2_j|^2
        ENSURE(m, B_XINV, (size_t)Ng * Ng * 8);
        ENSURE(m, B_SPG, (size_t)Nb * Ng * 8);
        ENSURE(m, B_SPH, (size_t)Nb * std::max(Ng, Nb) * 8);
        double* X1 = m->d(B_XINV);
        trtri_blocked(c, L1, m->d(B_DINV), Ng, X1);
        colsumsq_f64(s, X1, Ng, Ng, Ng, t1, 1);  // [B1^-1]_ii
        double* G = m->d(B_SPG);
        double* H = m->d(B_SPH);
        gemm2_f64(s, 1, 0, Nb, Ng, Ng, 1.0, aux, Ng, X1, Ng, 0.0, G, Nb, 0, 0);
        scale_rows_copy_f64(s, G, Nb, Nb, Ng, m->d(B_IV2), H);
        coldot_f64(s, G, H, Nb, Nb, Ng, t2);  // g^T V2^-1 g
        gemm2_f64(s, 1, 0, Nb, Ng, Nb, 1.0, Lk, Nb, H, Nb, 0.0, G, Nb, 0, 0);
        trsm_lower(s, Ac, Nb, m->d(B_DA), Nb, G, Nb, Ng);
        colsumsq_f64(s, G, Nb, Nb, Ng, t3, 0);  // |A^-1/2 Lk^T V2^-1 g|^2
        vec_axpby(s, Ng, 1.0, t2, -1.0, t3, t2);
        vec_axpby(s, Ng, 1.0, t1, 1.0, t2, t1);
        vec_mul(s, Ng, m->d(B_S1), t1, t1);
        vec_mul(s, Ng, m->d(B_S1), t1, t1);  // s_i^2 ([B1^-1]_ii + g^T S^-1 g)
        fill_f64(s, full, N, 0.0);
        scatter_vec_f64(s, t1, gi, ng, full);
        // bad sites: Q = A^-1/2 Lk^T V2^-1
        scale_rows_copy_f64(s, Lk, Nb, Nb, Nb, m->d(B_IV2), Tq);
        transpose_f64(s, Tq, Nb, Nb, Nb, H, Nb);
        trsm_lower(s, Ac, Nb, m->d(B_DA), Nb, H, Nb, Nb);
        colsumsq_f64(s, H, Nb, Nb, Nb, t2, 0);
        vec_axpby(s, Nb, 1.0, m->d(B_IV2), -1.0, t2, t2);
        scatter_vec_f64(s, t2, bi, nb, full);
        ep_split_var(s, n, m->d(B_V), full, m->d(B_FVAR));
    }
    double h[6];
    HIP_TRY(hipMemcpyAsync(h, sc + 10, 6 * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // h: [0] loglik_good / site terms, [1] f1^T a1 / z1^T a1, [2] sum log L1_ii,
    //    [3] 1/2 log|A|, [4] |Lk^-1 (f2 - m)|^2 / (m - z2)^T S^-1 (m - z2), [5] loglik_bad / bad terms
    const double le1 = h[0] - 0.5 * h[1] - h[2];
    const double le2 = ep ? (-0.5 * h[4] + h[3] + h[5]) : (h[5] - 0.5 * h[4] + h[3]);
    *logZ = le1 + le2;
    m->split = true;
    return GPL_OK;
}

// Predictive variance reduction k^T (K + V)^-1 k at M test columns for a split
// fit: KtT (N x M, rows in site order) -> red[M].  With q = L1^-1 S1 k_good and
// d = k_bad - aux^T q,  k^T (K + V)^-1 k = |q|^2 + d^T S^-1 d.
int gpl::split_var_reduction(gpl_ctx* c, gpl_model* m, const double* KtT, int64_t M, double* red) {
    hipStream_t s = c->stream;
    const int64_t N = m->N, Ng = m->Ng, Nb = m->Nb, ng = m->ng, nb = m->nb;
    const int64_t* gi = m->b[B_SIDX].as<int64_t>();
    const int64_t* bi = gi + ng;
    double* Q = (double*)c->buf(16, (size_t)Ng * M * 8);
    double* D = (double*)c->buf(17, (size_t)Nb * M * 8);
    double* E = (double*)c->buf(18, (size_t)Nb * M * 8);
    double* Hh = (double*)c->buf(19, (size_t)Nb * M * 8 + (size_t)M * 3 * 8);
    if (!Q || !D || !E || !Hh) return fail(GPL_ERR_OOM, "device allocation failed");
    double* r1 = Hh + Nb * M;
    double* r2 = r1 + M;
    gather_rows_f64(s, KtT, N, gi, ng, Ng, M, m->d(B_S1), Q, Ng);
    trsm_lower(s, m->d(B_L), Ng, m->d(B_DINV), Ng, Q, Ng, M);
    colsumsq_f64(s, Q, Ng, Ng, M, red, 0);
    gather_rows_f64(s, KtT, N, bi, nb, Nb, M, nullptr, D, Nb);
    gemm2_f64(s, 1, 0, Nb, M, Ng, -1.0, m->d(B_AUX), Ng, Q, Ng, 1.0, D, Nb, 0, 0);
    scale_rows_copy_f64(s, D, Nb, Nb, M, m->d(B_IV2), E);
    coldot_f64(s, D, E, Nb, Nb, M, r1);
    gemm2_f64(s, 1, 0, Nb, M, Nb, 1.0, m->d(B_LK), Nb, E, Nb, 0.0, Hh, Nb, 0, 0);
    trsm_lower(s, m->d(B_AC), Nb, m->d(B_DA), Nb, Hh, Nb, M);
    colsumsq_f64(s, Hh, Nb, Nb, M, r2, 0);
    vec_axpby(s, M, 1.0, r1, -1.0, r2, r1);
    vec_axpby(s, M, 1.0, red, 1.0, r1, red);
    return GPL_OK;
}
