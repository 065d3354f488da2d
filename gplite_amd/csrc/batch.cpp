// Batched form: energy + gradient of many independent Gaussian-likelihood GPs
// of uneven size in one device pass.  Problems are grouped into buckets by
// their padded size N (multiple of 64; the padding is an identity block, so
// every result is exactly that of the unpadded problem).  Each bucket runs the
// mixed-precision pipeline -- FP32 B build, FP32-MFMA blocked Cholesky with
// FP64 64x64 diagonal blocks and panel solves, FP32 L^-1 and B^-1, FP64
// iterative refinement of alpha with a matrix-free FP64 K*v, fused gradient --
// as batched launches (problem = grid row).  The whole multi-bucket launch
// sequence is captured once into a hipGraph and replayed.  The GEMM tile edge
// and the outer panel width per bucket come from gpl_batch_tune (timed
// candidates) or gpl_batch_set_tuning (a persisted table).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "host_common.h"
#include "model.h"

using namespace gpl;

namespace {

enum {
    BB_X, BB_XOFF, BB_N, BB_PROG, BB_F, BB_Y, BB_SV, BB_SCAL, BB_SIG, BB_WSC, BB_IDX, BB_RHS, BB_U, BB_R, BB_D, BB_T1,
    BB_KU, BB_ALPHA, BB_L, BB_X32, BB_BI, BB_DINV, BB_LDIAG, BB_INFO, BB_PART, BB_KMV, BB_RED, BB_CORR, BB_DACC, BB_B64, BB_RELMAX, BB_COUNT
};

struct Bucket {
    int64_t N = 0;
    std::vector<int> idx;  // problem indices
    int nfeat = 1;
    int bm = 0;       // GEMM tile edge (0 = heuristic)
    int nbo = 256;    // outer panel width
    int nchunk = 1;
    DevBuf b[BB_COUNT];
    double* d(int i) const { return b[i].as<double>(); }
    float* f(int i) const { return b[i].as<float>(); }
    ~Bucket() {
        for (auto& x : b) x.free_();
    }
};

// One stream pair per lane: buckets on different lanes run concurrently (their
// latency-bound diagonal chains overlap); inside a bucket the side stream
// carries the look-ahead panel and the refinement.
struct Lane {
    hipStream_t s = nullptr, side = nullptr;
    hipEvent_t ev_a = nullptr, ev_p = nullptr, done = nullptr;
};

}  // namespace

struct gpl_batch {
    gpl_ctx* ctx = nullptr;
    int nprob = 0, nterms = 0, ng = 0;
    gpl_opts opts{};
    DevProg hp{};  // term structure shared by all problems (host copy of problem 0)
    std::vector<Bucket*> buckets;
    std::vector<int> bucket_of;  // problem -> bucket index
    std::vector<int> pos_in;     // problem -> position in its bucket
    DevBuf res_logz, res_grad;
    std::vector<hipGraph_t> graph;      // one per lane
    std::vector<hipGraphExec_t> exec;   // one per lane (null: lane without buckets)
    std::vector<Lane> lanes;
    std::vector<int> lane_of;  // bucket -> lane
    std::vector<int32_t> status;       // per problem, from the last pass
    std::vector<double> refine_rel;    // per problem, last refinement correction of the last pass
    hipEvent_t fork = nullptr;
    std::vector<const double*> xs;  // gpl_fit_batch: the callers' inputs, for the exported models
    ~gpl_batch() {
        for (auto* b : buckets) delete b;
        res_logz.free_();
        res_grad.free_();
        for (auto e : exec)
            if (e) hipGraphExecDestroy(e);
        for (auto g : graph)
            if (g) hipGraphDestroy(g);
        for (auto& l : lanes) {
            if (l.s) hipStreamDestroy(l.s);
            if (l.side) hipStreamDestroy(l.side);
            if (l.ev_a) hipEventDestroy(l.ev_a);
            if (l.ev_p) hipEventDestroy(l.ev_p);
            if (l.done) hipEventDestroy(l.done);
        }
        if (fork) hipEventDestroy(fork);
    }
};

// A pass runs a fixed number of refinement sweeps.  From the last two corrections
// c1, c2 (relative to the solution) the contraction is rho = c2 / c1 and the error
// left after the last sweep about c2 rho / (1 - rho); it must be below this.  A
// last correction under 1e-11 is at the FP64 residual's rounding floor, where the
// corrections stop shrinking (rho ~ 1) and the solve is converged.
static const double kBatchRefineTol = 1e-8;

static std::string fmt_e(double v) {
    char b[32];
    snprintf(b, sizeof b, "%.3e", v);
    return b;
}

static double refine_error_estimate(double last, double prev) {
    if (last != last || prev != prev) return NAN;
    if (last <= 1e-3 * kBatchRefineTol) return last;
    const double rho = prev > 0.0 ? last / prev : 1.0;
    return rho < 1.0 ? last * rho / (1.0 - rho) : INFINITY;
}

// scalars per problem in BB_RED: five fixed slots and the gradient partials
static int red_stride(const gpl_batch* B) { return 6 + std::max(1, B->hp.ngrad); }

// --------------------------------------------------------------------------
// Recursive panel steps (every problem of the bucket at once).  A panel of W
// columns is split into halves: factor the left half, update the right half
// with one GEMM of depth W/2 over all rows below, factor the right half.  The
// 64-column leaves factor their FP64 diagonal tile and solve the rows below in
// FP64; the update GEMMs get depths 256/128/64 instead of a chain of depth-64
// updates.
// --------------------------------------------------------------------------
static void potrf_panel_rec(gpl_ctx* c, hipStream_t s, Bucket& k, int64_t K0, int64_t W) {
    const int nb = (int)k.idx.size();
    const int64_t N = k.N, N2 = N * N;
    float* L = k.f(BB_L);
    double* Dinv = k.d(BB_DINV);
    double* Dacc = k.d(BB_DACC);
    if (W == 64) {
        const int64_t kk = K0 / 64;
        {
            Prof p(c, "batch_potrf_diag", 0, 0, s);
            potrf_diag3_f32_batched(s, L, N, kk, Dinv, k.d(BB_LDIAG), k.b[BB_INFO].as<int>(), nb, N2, N * 64, N, Dacc);
        }
        const int64_t M = N - (K0 + 64);
        if (M > 0) {
            float* L21 = L + (K0 + 64) + K0 * N;
            Prof p(c, "batch_trsm_f64", 2.0 * nb * M * 64 * 64, 0, s);
            gemm_mixed_f32f64_batched(s, 0, 1, M, 64, 64, 1.0, L21, N, Dinv + kk * 64 * 64, 64, 0.0, L21, N, 0, nb, N2,
                                      N * 64, N2);
        }
        return;
    }
    const int64_t h = (W / 128) * 64, W2 = W - h, M = N - (K0 + h);
    potrf_panel_rec(c, s, k, K0, h);
    const float* P = L + (K0 + h) + K0 * N;
    {
        // k is at most half a panel here: 64-wide tiles (shorter prologue and
        // epilogue per k, four times the tiles) whatever the bucket's tile
        Prof p(c, "batch_panel_f32", 2.0 * nb * M * W2 * h, 0, s);
        gemm_f32_batched(s, 0, 1, M, W2, h, -1.f, P, N, P, N, 1.f, L + (K0 + h) * (N + 1), N, 0, 64, nb, N2, N2,
                         N2);
    }
    {
        Prof p(c, "batch_dacc_f64", 2.0 * nb * 4096.0 * h * (W2 / 64), 0, s);
        dacc_update(s, Dacc + ((K0 + h) / 64) * 64 * 64, P, N, W2 / 64, h, nb, N * 64, N2);
    }
    potrf_panel_rec(c, s, k, K0 + h, W2);
}

// X = L^-1 for every problem of the bucket by recursive doubling from the FP64
// diagonal-block inverses: at level b each pair of adjacent b-blocks
// [[X11, 0], [L21, X22]] gets X21 = -X22 (L21 X11).  One launch per product and
// level covers all pairs (grid rows) of all problems (grid z); T (the B^-1
// buffer, free until B^-1 is formed) holds L21 X11.
static void trtri_doubling_batched(gpl_ctx* c, hipStream_t s, Bucket& k) {
    const int nb = (int)k.idx.size();
    const int64_t N = k.N, N2 = N * N;
    const float* L = k.f(BB_L);
    float* X = k.f(BB_X32);
    float* T = k.f(BB_BI);
    {
        Prof p(c, "batch_trtri_leaves", 0, 0, s);
        trtri_leaves_f32_batched(s, X, N, k.d(BB_DINV), nb, N2, N * 64, 0, N);
    }
    for (int64_t b = 64; b < N; b *= 2) {
        const int64_t full = N / (2 * b);
        const int64_t r = std::min<int64_t>(b, N - full * 2 * b - b);
        const int64_t step = 2 * b * (N + 1);
        Prof p(c, "batch_trtri_level", 2.0 * nb * ((double)b * b * b * full + (r > 0 ? (double)r * b * (b + r) / 2 : 0.0)),
               0, s);
        if (full > 0) {
            gemm_f32_batched2(s, 0, 0, b, b, b, 1.f, L + b, N, X, N, 0.f, T + b, N, 4, 0, (int)full, step, step, step,
                              nb, N2, kFoldInverse);
            gemm_f32_batched2(s, 0, 0, b, b, b, -1.f, X + b * (N + 1), N, T + b, N, 0.f, X + b, N, 3, 0, (int)full,
                              step, step, step, nb, N2, kFoldInverse);
        }
        if (r > 0) {
            const int64_t o = full * step;
            gemm_f32_batched(s, 0, 0, r, b, b, 1.f, L + o + b, N, X + o, N, 0.f, T + o + b, N, 4, 0, nb, N2, N2, N2,
                             kFoldInverse);
            gemm_f32_batched(s, 0, 0, r, b, r, -1.f, X + o + b * (N + 1), N, T + o + b, N, 0.f, X + o + b, N, 3, 0,
                             nb, N2, N2, N2, kFoldInverse);
        }
    }
}

// --------------------------------------------------------------------------
// the per-bucket pipeline (enqueued on one stream; captured into the graph)
// --------------------------------------------------------------------------
static void bucket_pipeline(gpl_batch* B, Bucket& k, Lane& ln) {
    hipStream_t s = ln.s;
    gpl_ctx* c = B->ctx;
    const int nb = (int)k.idx.size();
    const int64_t N = k.N, N2 = N * N, T = N / 64;
    const int64_t sF = N * MAX_FEAT;
    const DevProg* progs = k.b[BB_PROG].as<DevProg>();
    const int64_t* nvec = k.b[BB_N].as<int64_t>();
    float* L = k.f(BB_L);
    float* X = k.f(BB_X32);
    float* Bi = k.f(BB_BI);
    double* Dacc = k.d(BB_DACC);  // FP64 diagonal tiles, same layout as Dinv
    double* ldiag = k.d(BB_LDIAG);
    const int bm = k.bm;
    const int64_t NBo = k.nbo;
    {
        Prof p(c, "batch_build_B_f32", 0, 4.0 * nb * N * (N + 64) / 2, s);
        launch_build_B_f32_batched(s, progs, k.nfeat, k.d(BB_F), sF, nvec, N, B->opts.jitter, k.d(BB_SCAL), L, N2, nb,
                                   Dacc, N * 64, k.d(BB_B64));
    }
    // blocked Cholesky (FP32 storage, FP64 diagonal blocks and panel solves) with
    // one panel of look-ahead: the next panel's columns are updated first and
    // factored on the side stream while the main stream applies the rest of the
    // trailing update.  The fork and join are events, so graph capture records
    // them as dependencies between the two streams.
    hipStream_t s2 = ln.side;
    potrf_panel_rec(c, s, k, 0, std::min<int64_t>(NBo, N));
    for (int64_t K0 = 0; K0 < N; K0 += NBo) {
        const int64_t W = std::min<int64_t>(NBo, N - K0);
        const int64_t K1 = K0 + W, R = N - K1;
        if (R <= 0) break;
        const int64_t W1 = std::min<int64_t>(NBo, R);
        const float* P = L + K1 + K0 * N;
        {
            Prof p(c, "batch_syrk_f32", 2.0 * nb * R * W1 * W, 0, s);
            gemm_f32_batched(s, 0, 1, R, W1, W, -1.f, P, N, P, N, 1.f, L + K1 * (N + 1), N, 0, bm, nb, N2, N2, N2);
        }
        {
            Prof p(c, "batch_dacc_f64", 2.0 * nb * 4096.0 * W * (W1 / 64), 0, s);
            dacc_update(s, Dacc + (K1 / 64) * 64 * 64, P, N, W1 / 64, W, nb, N * 64, N2);
        }
        hipEventRecord(ln.ev_a, s);
        hipStreamWaitEvent(s2, ln.ev_a, 0);
        potrf_panel_rec(c, s2, k, K1, W1);
        hipEventRecord(ln.ev_p, s2);
        const int64_t R2 = R - W1;
        if (R2 > 0) {
            const float* P2 = L + (K1 + W1) + K0 * N;
            {
                Prof p(c, "batch_syrk_f32", (double)nb * R2 * R2 * W, 0, s);
                gemm_f32_batched(s, 0, 1, R2, R2, W, -1.f, P2, N, P2, N, 1.f, L + (K1 + W1) * (N + 1), N, 1, bm, nb, N2,
                                 N2, N2);
            }
            {
                Prof p(c, "batch_dacc_f64", 2.0 * nb * 4096.0 * W * (R2 / 64), 0, s);
                dacc_update(s, Dacc + ((K1 + W1) / 64) * 64 * 64, P2, N, R2 / 64, W, nb, N * 64, N2);
            }
        }
        hipStreamWaitEvent(s, ln.ev_p, 0);
    }
    // X = L^-1 (FP32, FP64 diagonal-block inverses as leaves)
    trtri_doubling_batched(c, s, k);
    // alpha = S B^-1 S y by iterative refinement (fixed sweep count: the graph is
    // static) on the side stream, while the main stream forms B^-1 = X^T X
    hipEventRecord(ln.ev_a, s);
    hipStreamWaitEvent(s2, ln.ev_a, 0);
    const int64_t NV = (int64_t)nb * N;
    double* rhs = k.d(BB_RHS);
    double* u = k.d(BB_U);
    double* r = k.d(BB_R);
    double* dd = k.d(BB_D);
    double* t1 = k.d(BB_T1);
    double* Ku = k.d(BB_KU);
    double* sv = k.d(BB_SV);
    seg_scale(s2, nullptr, k.d(BB_SCAL), nvec, N, sv, nb);      // S = 1/sigma on valid rows, 0 on padding
    vec_mul(s2, NV, sv, k.d(BB_Y), rhs);                        // S y
    fill_f64(s2, u, NV, 0.0);
    vec_axpby(s2, NV, 1.0, rhs, 0.0, nullptr, r);
    const int sweeps = B->opts.refine_iters > 0 ? B->opts.refine_iters : 4;
    for (int it = 0; it < sweeps; ++it) {
        {
            Prof p(c, "batch_refine_f32", 0, 0, s2);
            trmv_lower_f32_batched(s2, X, N, N, r, t1, 0, nb, N2, N, k.d(BB_KMV), k.nchunk);
            trmv_lower_f32_batched(s2, X, N, N, t1, dd, 1, nb, N2, N, k.d(BB_KMV), k.nchunk);
        }
        vec_axpby(s2, NV, 1.0, dd, 1.0, u, u);
        // the last two corrections relative to the solution: their ratio is the contraction
        if (it == sweeps - 2) seg_relmax(s2, dd, u, nvec, N, k.d(BB_RELMAX) + nb, nb);
        if (it == sweeps - 1) seg_relmax(s2, dd, u, nvec, N, k.d(BB_RELMAX), nb);
        {  // r = S y - B u with the FP64 copy of B
            Prof p(c, "batch_symv_b64", 2.0 * nb * N * N, 0, s2);
            symv_lt_f64(s2, k.d(BB_B64), N, u, Ku, k.d(BB_KMV), k.nchunk, nb, N2, N);
        }
        vec_axpby(s2, NV, 1.0, rhs, -1.0, Ku, r);
    }
    double* alpha = k.d(BB_ALPHA);
    vec_mul(s2, NV, sv, u, alpha);
    hipEventRecord(ln.ev_p, s2);
    // B^-1 = X^T X and the gradient
    {
        Prof p(c, "batch_lauum_f32", (double)nb * N * N * N / 3, 0, s);
        gemm_f32_batched(s, 1, 0, N, N, N, 1.f, X, N, X, N, 0.f, Bi, N, 2, bm, nb, N2, N2, N2, kFoldInverse);
    }
    hipStreamWaitEvent(s, ln.ev_p, 0);
    const int64_t ntiles = T * (T + 1) / 2;
    {
        Prof p(c, "batch_grad", 0, 0, s);
        launch_grad_tiles_f32_batched(s, progs, B->hp, k.d(BB_F), sF, nvec, N, Bi, N2, k.d(BB_WSC), alpha, N,
                                      k.d(BB_PART), nb);
    }
    {
        Prof p(c, "batch_round_corr", 0, 0, s);
        launch_round_corr(s, k.d(BB_B64), Bi, N, k.d(BB_CORR), nb, N2, N2);
    }
    double* red = k.d(BB_RED);
    const int rs = red_stride(B);
    seg_dot(s, k.d(BB_Y), alpha, nvec, N, red, rs, 0, nb);
    seg_dot(s, ldiag, nullptr, nvec, N, red, rs, 1, nb);
    seg_dot(s, alpha, alpha, nvec, N, red, rs, 2, nb);
    colsumsq_f32_batched(s, X, N, N, N, t1, 1, nb, N2, N);
    seg_dot(s, t1, nullptr, nvec, N, red, rs, 3, nb);
    seg_cols(s, k.d(BB_CORR), ntiles, 1, red, rs, 4, nb);
    seg_cols(s, k.d(BB_PART), ntiles, B->hp.ngrad, red, rs, 5, nb);
    batch_finish(s, red, rs, nvec, k.d(BB_SIG), k.b[BB_IDX].as<int32_t>(), B->hp.ngrad, nb, B->res_logz.as<double>(),
                 B->res_grad.as<double>(), B->ng);
}

// Every bucket on its lane: the lanes fork from `root` and join back into it.
static void run_buckets(gpl_batch* B, hipStream_t root) {
    hipEventRecord(B->fork, root);
    std::vector<char> used(B->lanes.size(), 0);
    for (size_t i = 0; i < B->buckets.size(); ++i) {
        Lane& ln = B->lanes[B->lane_of[i]];
        if (!used[B->lane_of[i]]) {
            hipStreamWaitEvent(ln.s, B->fork, 0);
            used[B->lane_of[i]] = 1;
        }
        bucket_pipeline(B, *B->buckets[i], ln);
    }
    for (size_t l = 0; l < B->lanes.size(); ++l)
        if (used[l]) {
            hipEventRecord(B->lanes[l].done, B->lanes[l].s);
            hipStreamWaitEvent(root, B->lanes[l].done, 0);
        }
}

// one bucket on lane 0 (autotuner)
static void run_bucket(gpl_batch* B, Bucket& k, hipStream_t root) {
    Lane& ln = B->lanes[0];
    hipEventRecord(B->fork, root);
    hipStreamWaitEvent(ln.s, B->fork, 0);
    bucket_pipeline(B, k, ln);
    hipEventRecord(ln.done, ln.s);
    hipStreamWaitEvent(root, ln.done, 0);
}

static int upload_params(gpl_batch* B, const gpl_cf_term* terms, const double* sigma) {
    hipStream_t s = B->ctx->stream;
    for (auto* k : B->buckets) {
        const int nb = (int)k->idx.size();
        std::vector<DevProg> progs(nb);
        std::vector<double> scal(nb), sig(nb), wsc(nb);
        for (int q = 0; q < nb; ++q) {
            const int p = k->idx[q];
            int ncols = 0;
            for (int t = 0; t < B->nterms; ++t)
                ncols = std::max(ncols, terms[(size_t)p * B->nterms + t].col0 + terms[(size_t)p * B->nterms + t].ncols);
            int r = prepare_prog(terms + (size_t)p * B->nterms, B->nterms, ncols, progs[q]);
            if (r) return r;
            if (progs[q].ngrad != B->hp.ngrad || progs[q].t[0].kind != B->hp.t[0].kind ||
                progs[q].t[0].nl != B->hp.t[0].nl)
                return fail(GPL_ERR_ARG, "all problems of a batch need the same covariance structure");
            if (!(sigma[p] > 0)) return fail(GPL_ERR_ARG, "sigma must be positive");
            scal[q] = 1.0 / sigma[p];
            sig[q] = sigma[p];
            wsc[q] = 1.0 / (sigma[p] * sigma[p]);
        }
        HIP_TRY(hipMemcpyAsync(k->b[BB_PROG].p, progs.data(), nb * sizeof(DevProg), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(k->b[BB_SCAL].p, scal.data(), nb * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(k->b[BB_SIG].p, sig.data(), nb * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(k->b[BB_WSC].p, wsc.data(), nb * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    return GPL_OK;
}

static void drop_graph(gpl_batch* B) {
    for (auto e : B->exec)
        if (e) hipGraphExecDestroy(e);
    for (auto g : B->graph)
        if (g) hipGraphDestroy(g);
    B->exec.clear();
    B->graph.clear();
}

extern "C" int gpl_batch_create(gpl_ctx* c, int32_t nprob, const gpl_cf_term* terms, int32_t nterms,
                                const double* const* x, const int64_t* n, const int64_t* ncols,
                                const double* const* y, const double* sigma, const gpl_opts* opts, gpl_batch** out) {
    if (!c || !terms || !x || !n || !ncols || !y || !sigma || !out || nprob < 1)
        return fail(GPL_ERR_ARG, "NULL argument or empty batch");
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    gpl_batch* B = new gpl_batch();
    B->ctx = c;
    B->nprob = nprob;
    B->nterms = nterms;
    if (opts) B->opts = *opts;
    else gpl_opts_default(&B->opts);
    if (!opts) B->opts.refine_iters = 4;
    int r = prepare_prog(terms, nterms, ncols[0], B->hp);
    if (r) {
        delete B;
        return r;
    }
    B->ng = B->hp.ngrad + 1;
    {
        // kernels size their feature staging and the gradient fast path by the widest problem
        int maxc = 0;
        for (int p = 0; p < nprob; ++p) maxc = std::max<int64_t>(maxc, ncols[p]);
        DevProg widest;
        std::vector<gpl_cf_term> tw(terms + (size_t)0, terms + nterms);
        for (int p = 0; p < nprob; ++p)
            if (ncols[p] == maxc) {
                for (int t = 0; t < nterms; ++t) tw[t] = terms[(size_t)p * nterms + t];
                break;
            }
        r = prepare_prog(tw.data(), nterms, maxc, widest);
        if (r) {
            delete B;
            return r;
        }
        B->hp = widest;
    }
    // buckets by padded size
    std::map<int64_t, int> bidx;
    B->bucket_of.resize(nprob);
    B->pos_in.resize(nprob);
    for (int p = 0; p < nprob; ++p) {
        if (n[p] < 1) {
            delete B;
            return fail(GPL_ERR_ARG, "every problem needs n >= 1");
        }
        const int64_t Np = pad64(n[p]);
        auto it = bidx.find(Np);
        if (it == bidx.end()) {
            bidx[Np] = (int)B->buckets.size();
            Bucket* k = new Bucket();
            k->N = Np;
            B->buckets.push_back(k);
            it = bidx.find(Np);
        }
        Bucket* k = B->buckets[it->second];
        B->bucket_of[p] = it->second;
        B->pos_in[p] = (int)k->idx.size();
        k->idx.push_back(p);
    }
    hipStream_t s = c->stream;
    for (auto* k : B->buckets) {
        const int nb = (int)k->idx.size();
        const int64_t N = k->N, T = N / 64;
        k->nchunk = kmatvec_chunks(N);
        k->nfeat = std::max(1, B->hp.nfeat);
        std::vector<int64_t> xoff(nb), nv(nb);
        std::vector<int32_t> idx(nb);
        int64_t tot = 0;
        for (int q = 0; q < nb; ++q) {
            const int p = k->idx[q];
            xoff[q] = tot;
            nv[q] = n[p];
            idx[q] = p;
            tot += n[p] * ncols[p];
        }
        const size_t vec = (size_t)nb * N * 8, mat32 = (size_t)nb * N * N * 4;
        bool ok = k->b[BB_X].ensure((size_t)std::max<int64_t>(tot, 1) * 8) && k->b[BB_XOFF].ensure(nb * 8) &&
                  k->b[BB_N].ensure(nb * 8) && k->b[BB_PROG].ensure(nb * sizeof(DevProg)) &&
                  k->b[BB_F].ensure((size_t)nb * N * MAX_FEAT * 8) && k->b[BB_SCAL].ensure(nb * 8) &&
                  k->b[BB_SIG].ensure(nb * 8) && k->b[BB_WSC].ensure(nb * 8) && k->b[BB_IDX].ensure(nb * 4) &&
                  k->b[BB_L].ensure(mat32) && k->b[BB_X32].ensure(mat32) && k->b[BB_BI].ensure(mat32) &&
                  k->b[BB_B64].ensure(mat32 * 2) && k->b[BB_RELMAX].ensure((size_t)nb * 16) &&
                  k->b[BB_DINV].ensure((size_t)nb * N * 64 * 8) && k->b[BB_DACC].ensure((size_t)nb * N * 64 * 8) &&
                  k->b[BB_LDIAG].ensure(vec) &&
                  k->b[BB_INFO].ensure(nb * 4 + 64) && k->b[BB_PART].ensure((size_t)nb * T * (T + 1) / 2 * 8 * std::max(1, B->hp.ngrad)) &&
                  k->b[BB_KMV].ensure((size_t)nb * std::max(k->nchunk, symv_slots(N, nb)) * N * 8) &&
                  k->b[BB_RED].ensure((size_t)nb * red_stride(B) * 8) &&
                  k->b[BB_CORR].ensure((size_t)nb * T * (T + 1) / 2 * 8);
        for (int i : {BB_Y, BB_SV, BB_RHS, BB_U, BB_R, BB_D, BB_T1, BB_KU, BB_ALPHA}) ok = ok && k->b[i].ensure(vec);
        if (!ok) {
            delete B;
            return fail(GPL_ERR_OOM, "device allocation failed for the batch");
        }
        // concatenated raw inputs and padded targets
        std::vector<double> xs((size_t)std::max<int64_t>(tot, 1)), ys((size_t)nb * N, 0.0);
        for (int q = 0; q < nb; ++q) {
            const int p = k->idx[q];
            memcpy(xs.data() + xoff[q], x[p], n[p] * ncols[p] * 8);
            memcpy(ys.data() + (size_t)q * N, y[p], n[p] * 8);
        }
        HIP_TRY(hipMemcpyAsync(k->b[BB_X].p, xs.data(), xs.size() * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(k->b[BB_XOFF].p, xoff.data(), nb * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(k->b[BB_N].p, nv.data(), nb * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(k->b[BB_IDX].p, idx.data(), nb * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(k->b[BB_Y].p, ys.data(), ys.size() * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    if (!B->res_logz.ensure((size_t)nprob * 8) || !B->res_grad.ensure((size_t)nprob * B->ng * 8)) {
        delete B;
        return fail(GPL_ERR_OOM, "device allocation failed for the batch");
    }
    // lanes: up to sixteen stream pairs; buckets go to the least-loaded lane,
    // largest estimated time first (GEMM work plus a fixed cost per 64-row
    // diagonal step).  Most buckets are chains of small launches, so more lanes
    // overlap more of them: 2048 problems in 32 buckets take 99 ms per pass on 2
    // lanes, 86 on 4, 79 on 8, 76 on 16 (HIP multiplexes the streams onto its
    // hardware queues).
    {
        const int nl = (int)std::min<size_t>(16, B->buckets.size());
        B->lanes.resize(nl);
        hipError_t e = hipEventCreateWithFlags(&B->fork, hipEventDisableTiming);
        for (auto& l : B->lanes) {
            if (e == hipSuccess) e = hipStreamCreateWithFlags(&l.s, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipStreamCreateWithFlags(&l.side, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&l.ev_a, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&l.ev_p, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&l.done, hipEventDisableTiming);
        }
        if (e != hipSuccess) {
            delete B;
            return fail(GPL_ERR_HIP, std::string("lane streams: ") + hipGetErrorString(e));
        }
        std::vector<int> order(B->buckets.size());
        std::vector<double> cost(B->buckets.size()), load(nl, 0.0);
        for (size_t i = 0; i < order.size(); ++i) {
            const double N = (double)B->buckets[i]->N, nbk = (double)B->buckets[i]->idx.size();
            cost[i] = nbk * N * N * N / 60e9 + (N / 64) * 0.25;  // ms
            order[i] = (int)i;
        }
        std::sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
        B->lane_of.assign(B->buckets.size(), 0);
        for (int i : order) {
            const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            B->lane_of[i] = l;
            load[l] += cost[i];
        }
    }
    r = upload_params(B, terms, sigma);
    if (r) {
        delete B;
        return r;
    }
    // features (O(n d), once; hyperparameter updates do not change them for stationary terms)
    for (auto* k : B->buckets) {
        fill_f64(s, k->d(BB_F), (int64_t)k->idx.size() * k->N * MAX_FEAT, 0.0);
        launch_features_batched(s, k->d(BB_X), k->b[BB_XOFF].as<int64_t>(), k->b[BB_N].as<int64_t>(),
                                k->b[BB_PROG].as<DevProg>(), k->d(BB_F), k->N, k->N * MAX_FEAT, (int)k->idx.size());
    }
    HIP_TRY(hipStreamSynchronize(s));
    *out = B;
    return GPL_OK;
}

extern "C" int gpl_batch_set_params(gpl_batch* B, const gpl_cf_term* terms, const double* sigma) {
    if (!B || !terms || !sigma) return fail(GPL_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(B->ctx->mu);
    HIP_TRY(hipSetDevice(B->ctx->device));
    bool periodic = false;
    for (int p = 0; p < B->nprob; ++p)
        for (int t = 0; t < B->nterms; ++t) periodic = periodic || terms[(size_t)p * B->nterms + t].kind == GPL_CF_PERIODIC;
    int r = upload_params(B, terms, sigma);
    if (r || !periodic) return r;
    // periodic features [sin(w x), cos(w x)] and their period derivatives depend on
    // the period: rebuild them from the resident inputs (O(n d), on the context's
    // stream, ahead of the next pass)
    hipStream_t s = B->ctx->stream;
    for (auto* k : B->buckets) {
        fill_f64(s, k->d(BB_F), (int64_t)k->idx.size() * k->N * MAX_FEAT, 0.0);
        launch_features_batched(s, k->d(BB_X), k->b[BB_XOFF].as<int64_t>(), k->b[BB_N].as<int64_t>(),
                                k->b[BB_PROG].as<DevProg>(), k->d(BB_F), k->N, k->N * MAX_FEAT, (int)k->idx.size());
    }
    return GPL_OK;
}

static void reset_flags(gpl_batch* B, hipStream_t s) {
    // outside any graph: status words are cleared with plain memsets before a pass
    for (auto* k : B->buckets) hipMemsetAsync(k->b[BB_RELMAX].p, 0, k->idx.size() * 16, s);
    for (auto* k : B->buckets) hipMemsetAsync(k->b[BB_INFO].p, 0, k->idx.size() * 4, s);
}

// One graph per lane, captured on the lane's stream (its look-ahead side stream
// joins through events) and launched on it: the lanes' graphs then run
// concurrently, where a single graph holding every lane executed them one after
// the other (measured: 2048 uneven problems in 87.7 ms against 100.8 ms).
static int run_once(gpl_batch* B) {
    hipStream_t s = B->ctx->stream;
    reset_flags(B, s);
    const size_t nl = B->lanes.size();
    if (B->exec.empty()) {
        B->graph.assign(nl, nullptr);
        B->exec.assign(nl, nullptr);
        const bool prof = B->ctx->prof;
        B->ctx->prof = false;  // no event records inside the graphs
        for (size_t l = 0; l < nl; ++l) {
            Lane& ln = B->lanes[l];
            bool any = false;
            for (size_t i = 0; i < B->buckets.size(); ++i) any = any || B->lane_of[i] == (int)l;
            if (!any) continue;
            hipError_t e = hipStreamBeginCapture(ln.s, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                for (size_t i = 0; i < B->buckets.size(); ++i)
                    if (B->lane_of[i] == (int)l) bucket_pipeline(B, *B->buckets[i], ln);
                e = hipStreamEndCapture(ln.s, &B->graph[l]);
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&B->exec[l], B->graph[l], nullptr, nullptr, 0);
            if (e != hipSuccess) {
                B->ctx->prof = prof;
                drop_graph(B);
                return fail(GPL_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            }
        }
        B->ctx->prof = prof;
    }
    HIP_TRY(hipEventRecord(B->fork, s));
    for (size_t l = 0; l < nl; ++l) {
        if (!B->exec[l]) continue;
        Lane& ln = B->lanes[l];
        HIP_TRY(hipStreamWaitEvent(ln.s, B->fork, 0));
        HIP_TRY(hipGraphLaunch(B->exec[l], ln.s));
        HIP_TRY(hipEventRecord(ln.done, ln.s));
        HIP_TRY(hipStreamWaitEvent(s, ln.done, 0));
    }
    return GPL_OK;
}

// One pass (the graph, or eager launches when `eager` or profiling), then the
// results and the per-problem statuses.  Caller holds the context's lock.
static int batch_pass(gpl_batch* B, bool eager, double* log_evidence, double* grad) {
    Range rg("gpl_batch_pass");
    HIP_TRY(hipSetDevice(B->ctx->device));
    hipStream_t s = B->ctx->stream;
    if (eager || B->ctx->prof) {
        // profiled pass: eager launches with per-class events (no graph)
        reset_flags(B, s);
        run_buckets(B, s);
    } else {
        int r = run_once(B);
        if (r) return r;
    }
    if (log_evidence)
        HIP_TRY(hipMemcpyAsync(log_evidence, B->res_logz.p, (size_t)B->nprob * 8, hipMemcpyDeviceToHost, s));
    if (grad) HIP_TRY(hipMemcpyAsync(grad, B->res_grad.p, (size_t)B->nprob * B->ng * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // Non-PD pivots, and refinements whose last correction stays above 1e-8 of the
    // solution (cond(B) too large for the FP32 inverse factor with the pass's fixed
    // sweep count): those problems get log_evidence = -Inf and NaN gradients, the
    // others keep their results, and the call reports the first failing problem
    // (non-PD first).  gpl_batch_status gives every problem's status.
    int first = -1, first_nc = -1;
    int64_t first_minor = 0;
    B->status.assign(B->nprob, GPL_OK);
    B->refine_rel.assign(B->nprob, 0.0);
    for (auto* k : B->buckets) {
        const size_t nb = k->idx.size();
        std::vector<int> info(nb);
        std::vector<double> rr(2 * nb);
        HIP_TRY(hipMemcpy(info.data(), k->b[BB_INFO].p, nb * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(rr.data(), k->b[BB_RELMAX].p, nb * 16, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < nb; ++q) {
            const int pidx = k->idx[q];
            B->refine_rel[pidx] = rr[q];
            const bool notpd = info[q] > 0;
            const bool noconv = !(refine_error_estimate(rr[q], rr[nb + q]) <= kBatchRefineTol);
            if (!notpd && !noconv) continue;
            B->status[pidx] = notpd ? GPL_ERR_NOT_PD : GPL_ERR_NOCONV;
            if (log_evidence) log_evidence[pidx] = -INFINITY;
            if (grad)
                for (int j = 0; j < B->ng; ++j) grad[(size_t)pidx * B->ng + j] = NAN;
            if (notpd && (first < 0 || pidx < first)) {
                first = pidx;
                first_minor = info[q];
            }
            if (!notpd && (first_nc < 0 || pidx < first_nc)) first_nc = pidx;
        }
    }
    if (first >= 0) {
        set_minor(first_minor);
        return fail(GPL_ERR_NOT_PD, "problem " + std::to_string(first) + ": leading minor of order " +
                                        std::to_string(first_minor) + " is not positive");
    }
    if (first_nc >= 0)
        return fail(GPL_ERR_NOCONV, "problem " + std::to_string(first_nc) +
                                        ": iterative refinement did not converge to 1e-8 (last relative "
                                        "correction " + fmt_e(B->refine_rel[first_nc]) +
                                        "); refit it in FP64 or raise refine_iters");
    return GPL_OK;
}

extern "C" int gpl_batch_energy_grad(gpl_batch* B, double* log_evidence, double* grad) {
    if (!B) return fail(GPL_ERR_ARG, "NULL batch");
    std::lock_guard<std::mutex> lk(B->ctx->mu);
    return batch_pass(B, false, log_evidence, grad);
}

// Copies problem p's fit out of its bucket into the model (allocated or recycled).
static int export_model(gpl_batch* B, int p, const gpl_cf_term* terms, int64_t n, int64_t ncols, const double* y,
                        double sigma, double logz, gpl_model* m) {
    gpl_ctx* c = B->ctx;
    hipStream_t s = c->stream;
    const Bucket& k = *B->buckets[B->bucket_of[p]];
    const int64_t q = B->pos_in[p], N = k.N, N2 = N * N;
    DevProg P;
    int r = prepare_prog(terms + (size_t)p * B->nterms, B->nterms, ncols, P);
    if (r) return r;
    r = model_alloc(m, n, ncols, true);
    if (r) return r;
    ENSURE(m, B_XINV, (size_t)N2 * 4);
    m->prog = P;
    m->lik = gpl_lik{GPL_LIK_GAUSSIAN, GPL_LINK_IDENTITY, sigma, 0.0};
    m->opts = B->opts;
    m->opts.precision = GPL_PREC_MIXED;
    m->opts.approx = GPL_APPROX_LAPLACE;
    m->has_trials = m->has_offset = false;
    m->have_xinv = true;
    m->have_binv = false;
    m->split = m->no_posterior = m->refit_fp64 = false;
    m->logZ = logz;
    m->iters = 1;
    m->refine_rel = B->refine_rel[p];
    m->refine_iters_done = B->opts.refine_iters;
    auto cp = [&](int dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(m->b[dst].p, src, bytes, hipMemcpyDeviceToDevice, s);
    };
    HIP_TRY(cp(B_PROG, k.b[BB_PROG].as<DevProg>() + q, sizeof(DevProg)));
    HIP_TRY(cp(B_F, k.d(BB_F) + q * N * MAX_FEAT, (size_t)N * MAX_FEAT * 8));
    HIP_TRY(cp(B_L, k.f(BB_L) + q * N2, (size_t)N2 * 4));
    HIP_TRY(cp(B_XINV, k.f(BB_X32) + q * N2, (size_t)N2 * 4));
    HIP_TRY(cp(B_DINV, k.d(BB_DINV) + q * N * 64, (size_t)N * 64 * 8));
    HIP_TRY(cp(B_LDIAG, k.d(BB_LDIAG) + q * N, (size_t)N * 8));
    HIP_TRY(cp(B_SV, k.d(BB_SV) + q * N, (size_t)N * 8));
    HIP_TRY(cp(B_ALPHA, k.d(BB_ALPHA) + q * N, (size_t)N * 8));
    HIP_TRY(cp(B_Y, k.d(BB_Y) + q * N, (size_t)N * 8));
    HIP_TRY(cp(B_Z, k.d(BB_Y) + q * N, (size_t)N * 8));  // Gaussian pseudo-data z = y (R/lik.R:188-192)
    if (n * ncols > 0) HIP_TRY(hipMemcpyAsync(m->d(B_X), B->xs[p], n * ncols * 8, hipMemcpyHostToDevice, s));
    for (int i : {B_TR, B_OFF, B_FVAR, B_CAVM, B_CAVV}) fill_f64(s, m->d(i), N, 0.0);
    fill_f64(s, m->d(B_V), N, 0.0);
    fill_f64(s, m->d(B_V), n, sigma * sigma);
    // f = (K + jitter I) alpha = sigma (B u - u), B u from the last refinement sweep
    fill_f64(s, m->d(B_FMEAN), N, 0.0);
    vec_axpby(s, n, sigma, k.d(BB_KU) + q * N, -sigma, k.d(BB_U) + q * N, m->d(B_FMEAN));
    (void)y;
    return GPL_OK;
}

extern "C" int gpl_fit_batch(gpl_ctx* c, int32_t nprob, const gpl_cf_term* terms, int32_t nterms,
                             const double* const* x, const int64_t* n, const int64_t* ncols, const double* const* y,
                             const double* sigma, const gpl_opts* opts, gpl_model** models, double* log_evidence,
                             int32_t* status) {
    if (!c || !models || !status) return fail(GPL_ERR_ARG, "NULL argument");
    for (int p = 0; p < nprob; ++p)
        if (models[p] && model_ctx(models[p]) != c)
            return fail(GPL_ERR_ARG, "a model to refit belongs to another context");
    gpl_opts o;
    if (opts) o = *opts;
    else {
        gpl_opts_default(&o);
        o.refine_iters = 6;
    }
    gpl_batch* B = nullptr;
    int r = gpl_batch_create(c, nprob, terms, nterms, x, n, ncols, y, sigma, &o, &B);
    if (r) return r;
    B->xs.assign(x, x + nprob);
    // GEMM tiles by problem size alone, so that a problem's result does not depend
    // on the other problems of the call
    for (auto* k : B->buckets) k->bm = k->N >= 1024 ? 128 : 64;
    std::vector<double> lz(nprob);
    {
        std::lock_guard<std::mutex> lk(c->mu);
        r = batch_pass(B, true, lz.data(), nullptr);
        if (r == GPL_OK || r == GPL_ERR_NOT_PD || r == GPL_ERR_NOCONV) {
            r = GPL_OK;
            for (int p = 0; p < nprob && !r; ++p) {
                status[p] = B->status[p];
                if (log_evidence) log_evidence[p] = lz[p];
                if (status[p] != GPL_OK) continue;
                gpl_model* m = models[p];
                const bool fresh = !m;
                if (fresh) m = model_new(c);
                r = export_model(B, p, terms, n[p], ncols[p], y[p], sigma[p], lz[p], m);
                if (r && fresh) gpl_model_free(m);
                else models[p] = m;
            }
            if (r == GPL_OK) HIP_TRY(hipStreamSynchronize(c->stream));
        }
    }
    gpl_batch_free(B);
    return r;
}

extern "C" int gpl_batch_status(gpl_batch* B, int32_t* status, double* refine_rel) {
    if (!B) return fail(GPL_ERR_ARG, "NULL batch");
    if (B->status.empty()) return fail(GPL_ERR_ARG, "no pass has run yet");
    if (status) memcpy(status, B->status.data(), B->status.size() * 4);
    if (refine_rel) memcpy(refine_rel, B->refine_rel.data(), B->refine_rel.size() * 8);
    return GPL_OK;
}

extern "C" int32_t gpl_batch_grad_len(gpl_batch* B) { return B ? B->ng : 0; }

extern "C" int gpl_batch_info(gpl_batch* B, int32_t* nbuckets, double* refine_rel) {
    if (!B) return fail(GPL_ERR_ARG, "NULL batch");
    if (nbuckets) *nbuckets = (int32_t)B->buckets.size();
    if (refine_rel) {
        double mx = 0.0;
        for (double v : B->refine_rel) mx = (v != v || mx != mx) ? NAN : std::max(mx, v);
        *refine_rel = mx;
    }
    return GPL_OK;
}

extern "C" int gpl_batch_tune(gpl_batch* B, int32_t cap, int32_t* bucket_n, int32_t* bm, int32_t* nbo,
                              double* best_ms, int32_t* nbuckets) {
    if (!B) return fail(GPL_ERR_ARG, "NULL batch");
    std::lock_guard<std::mutex> lk(B->ctx->mu);
    HIP_TRY(hipSetDevice(B->ctx->device));
    hipStream_t s = B->ctx->stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const bool prof = B->ctx->prof;
    B->ctx->prof = false;
    int i = 0;
    for (auto* k : B->buckets) {
        float best = 1e30f;
        int bbm = 64, bnb = 256;
        for (int cbm : {64, 128})
            for (int cnb : {128, 256, 512, 1024}) {
                if (cnb > k->N && cnb != 128) continue;
                k->bm = cbm;
                k->nbo = cnb;
                run_bucket(B, *k, s);  // warm
                hipEventRecord(e0, s);
                run_bucket(B, *k, s);
                hipEventRecord(e1, s);
                hipEventSynchronize(e1);
                float ms = 0;
                hipEventElapsedTime(&ms, e0, e1);
                if (ms < best) best = ms, bbm = cbm, bnb = cnb;
            }
        k->bm = bbm;
        k->nbo = bnb;
        if (i < cap) {
            if (bucket_n) bucket_n[i] = (int32_t)k->N;
            if (bm) bm[i] = bbm;
            if (nbo) nbo[i] = bnb;
            if (best_ms) best_ms[i] = best;
        }
        ++i;
    }
    B->ctx->prof = prof;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    drop_graph(B);
    if (nbuckets) *nbuckets = (int32_t)B->buckets.size();
    return GPL_OK;
}

extern "C" int gpl_batch_set_tuning(gpl_batch* B, int32_t count, const int32_t* bucket_n, const int32_t* bm,
                                    const int32_t* nbo) {
    if (!B || (count > 0 && (!bucket_n || !bm || !nbo))) return fail(GPL_ERR_ARG, "NULL argument");
    for (int j = 0; j < count; ++j)
        for (auto* k : B->buckets)
            if (k->N == bucket_n[j]) {
                if ((bm[j] != 64 && bm[j] != 128) ||
                    (nbo[j] != 128 && nbo[j] != 256 && nbo[j] != 512 && nbo[j] != 1024))
                    return fail(GPL_ERR_ARG, "tile edge must be 64/128 and panel width 128/256/512/1024");
                k->bm = bm[j];
                k->nbo = nbo[j];
            }
    drop_graph(B);
    return GPL_OK;
}

extern "C" void gpl_batch_free(gpl_batch* B) {
    if (!B) return;
    hipSetDevice(B->ctx->device);
    hipStreamSynchronize(B->ctx->stream);
    delete B;
}
