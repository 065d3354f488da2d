// Internal declarations shared by the gfx950 kernels and the host orchestration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gplite_amd.h"

namespace gpl {

constexpr int TB = 64;              // tile edge: every dense matrix is padded to a multiple of 64
constexpr int MAX_FEAT = 64;        // device feature columns per model
constexpr int GRAD_MAX = 96;        // gradient entries computed on the device

// A covariance term after host preparation: features instead of raw columns.
//   stationary / lin / nn: features = the term's input columns
//   periodic: features = [sin(w x), cos(w x)] (2*ncols), plus 2*ncols derivative
//             features d/dlog(period) stored at df0
struct DevTerm {
    int kind, op;
    int c0, nc;        // raw input columns
    int f0, nf;        // feature columns
    int nl;            // number of length-scales (1 or nf)
    int base_kind;     // periodic base kernel
    int df0;           // periodic: derivative features
    int g0, ng;        // gradient slots [g0, g0+ng)
    int pad;
    double m2;   // magn^2
    double b0;   // nn: sigma0^2 sigma^2
    double s2;   // nn: sigma^2
    double w;    // periodic: 2 pi / period
    double il2[2 * GPL_MAX_COLS];  // 1/lscale^2 per feature column
};

struct DevProg {
    int nterms;
    int nfeat;     // feature columns in use
    int ngrad;     // gradient entries of the cf part
    int pad;
    DevTerm t[GPL_MAX_TERMS];
};

// ---- launchers (defined in the .hip files) -------------------------------
// Covariance (cov.hip)
void launch_features(hipStream_t s, const double* x, int64_t n, int64_t ldx, int ncols,
                     const DevProg* prog_dev, const DevProg& prog, double* F, int64_t ldf);
// out[i + j*ld] for the lower tiles of an N x N matrix built from F (N rows).
// mode 0: K + jitter*I; mode 1: B = I + s_i (K + jitter I)_ij s_j with s scalar
// (svec == null) or per-point; padded rows/cols (>= n) become identity.
void launch_build_sym_f64(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F,
                          int64_t n, int64_t N, double jitter, int mode, double sscal,
                          const double* svec, double* out, int64_t ld);
void launch_build_sym_f32(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F, int64_t n,
                          int64_t N, double jitter, int mode, double sscal, const double* svec, float* out,
                          int64_t ld, double* dacc, double* out64);
// out (Na x Nb col-major, ld) = K(Fa, Fb); zero outside [na) x [nb).
void launch_build_cross(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* Fa,
                        int64_t lda_f, int64_t na, int64_t Na, const double* Fb, int64_t ldb_f,
                        int64_t nb, int64_t Nb, double* out, int64_t ld);
// out[i] = k(Fa_i, Fb_i), i < n
void launch_build_diag(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* Fa,
                       int64_t lda_f, const double* Fb, int64_t ldb_f, int64_t n, double* out);
// y = (K + jitter I) v  (matrix-free FP64, recomputes K from F). part: N*nchunk scratch.
void launch_kmatvec(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F,
                    int64_t n, int64_t N, double jitter, const double* v, double* y, double* part,
                    int nchunk);
int kmatvec_chunks(int64_t N);
// gradient partials: part[tile * ng + p]; W is (K+V)^-1 given as c * Winv (lower tiles).
void launch_grad_tiles(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F,
                       int64_t n, int64_t N, const double* W, int64_t ldw, double wscale,
                       const double* alpha, double* part);
void launch_grad_tiles_f32(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F,
                           int64_t n, int64_t N, const float* W, int64_t ldw, double wscale,
                           const double* alpha, double* part);

// Dense FP64 (dense.hip)
// C[M x N] = beta*C + alpha * op(A) op(B); all dims multiples of 64.
// ta/tb: operand stored transposed.  tiles: 0 = full grid, 1 = lower tiles of a square C,
// 2 = lower tiles and k starts at the row tile (C = X^T X with X lower triangular),
// 3 = full grid and k stops after the row tile (C = X B with X lower triangular),
// 4 = full grid and k starts at the column tile (C = A X with X lower triangular).
// bm: output tile edge (64 or 128; 0 = choose by grid size).
void gemm2_f64(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
               int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int tiles, int bm);
void gemm2_f64_batched(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
                       int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int tiles,
                       int batch, int64_t sA, int64_t sB, int64_t sC);
// out[i] = sum_p W[p * stride + i] for i < n (fixed order)
void sum_parts_f64(hipStream_t s, const double* W, int64_t stride, int parts, int64_t n, double* out);
// FP64 MFMA on FP32 storage: C(f32) = beta*C + alpha op(A(f32)) op(B(f64)) and the mirror case.
void gemm_mixed_f32f64(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha, const float* A,
                       int64_t lda, const double* B, int64_t ldb, double beta, float* C, int64_t ldc, int tiles);
void gemm_mixed_f64f32(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                       const double* A, int64_t lda, const float* B, int64_t ldb, double beta, float* C, int64_t ldc,
                       int tiles);
void gemm_mixed_f32f64f64(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                          const float* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                          int64_t ldc, int tiles);
// FP32 storage and FP32 MFMA with FP32 accumulation (dense32.hip)
void gemm_f32(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, float alpha, const float* A,
              int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int tiles, int bm,
              int fold = 1);
void set_identity_f32(hipStream_t s, float* A, int64_t ld, int64_t N);
void trmv_lower_f32(hipStream_t s, const float* X, int64_t ld, int64_t N, const double* b, double* y, int trans,
                    double* part, int nchunk);
void colsumsq_f32(hipStream_t s, const float* A, int64_t ld, int64_t nrow, int64_t ncol, double* out, int lower);
void rel_maxabs(hipStream_t s, const double* d, const double* u, int64_t n, double* out);
// Factor the 64x64 diagonal block at A + k*64*(ld+1) in place (FP64 arithmetic); write its
// inverse to Dinv + k*64*64 (FP64) and log(L_ii) to ldiag + k*64; info = first failing row + 1.
void potrf_diag3_f64(hipStream_t s, double* A, int64_t ld, int64_t k, double* Dinv, double* ldiag, int* info,
                     int want_inv);
void potrf_diag3_f32(hipStream_t s, float* A, int64_t ld, int64_t k, double* Dinv, double* ldiag, int* info,
                     int want_inv, const double* Dacc);
// D_j -= P_j P_j^T in FP64 for ntiles consecutive diagonal tiles (FP32 panel P)
void dacc_update(hipStream_t s, double* D, const float* P, int64_t ldp, int64_t ntiles, int64_t kdepth, int batch,
                 int64_t sD, int64_t sP);
// triangular solve with the blocked factor (L lower, Dinv diagonal inverses), in place on b.
void trsv_lower_f64(hipStream_t s, const double* L, int64_t ld, const double* Dinv, int64_t N,
                    double* b, double* x, int trans);
// y = L^T? x with a dense lower-triangular X (GEMV on lower tiles): trans=0: y = X b; 1: y = X^T b
void trmv_lower_f64(hipStream_t s, const double* X, int64_t ld, int64_t N, const double* b,
                    double* y, int trans, double* part);
// y = X b or X^T b over nchunk column chunks of an FP64 lower-triangular X
// (partials in part: nchunk * N, summed in a fixed order; transposed tiles
// staged through LDS so every read is coalesced)
void trmv_lower_f64_chunked(hipStream_t s, const double* X, int64_t ld, int64_t N, const double* b, double* y,
                            int trans, double* part, int nchunk);
// y = K v for a symmetric K stored in its lower tiles.
void symv_lower_f64(hipStream_t s, const double* K, int64_t ld, int64_t N, const double* v,
                    double* y, double* part);
// out[j] = sum_i A[i + j*ld] * v[i] for j < ncol (one wavefront per column)
void gemv_t_f64(hipStream_t s, const double* A, int64_t ld, int64_t nrow, int64_t ncol, const double* v,
                double* out);
// out[j] = sum_i A[i + j*ld]^2 for j < ncol (deterministic column sums of squares)
void colsumsq_f64(hipStream_t s, const double* A, int64_t ld, int64_t nrow, int64_t ncol, double* out,
                  int lower);
// y[b][i] = sum over c of part[b][c][i] (fixed order)
void chunk_sum_f64(hipStream_t s, const double* part, int64_t N, int nchunk, double* y, int batch, int64_t sV);
// y = A v, A symmetric FP64 held as lower 64x64 tiles, each read once (part:
// batch * symv_slots(N, batch) * N doubles; nchunk is not used)
int symv_slots(int64_t N, int batch);
void symv_lt_f64(hipStream_t s, const double* A, int64_t N, const double* v, double* y, double* part, int nchunk,
                 int batch, int64_t sA, int64_t sV);
// W = s B^-1 s - (t u^T + u t^T) in place over the lower tiles of B^-1 (t, u may be null)
void form_wgrad_f64(hipStream_t s, double* B, int64_t N, int64_t n, const double* sv, const double* t,
                    const double* u);
void set_identity_f64(hipStream_t s, double* A, int64_t ld, int64_t N);
// info value a panel factorisation stores when one of its device waits timed out
constexpr int kPanelWaitTimeout = 0x7fffffff;
// the single-launch panel: at most this many 64-column blocks, and its flag words
constexpr int kPanelMaxBlocks = 16;
constexpr int kPanelSyncWords = kPanelMaxBlocks + kPanelMaxBlocks * kPanelMaxBlocks + 1;
// one outer panel (nc <= kPanelMaxBlocks column blocks from column K0, nrb = (ld - K0) / 64
// row blocks) of the FP64 blocked Cholesky in one launch; sync: kPanelSyncWords words
// of the context (zero at creation), epoch: nonzero and new for every launch, base:
// tickets handed out before (nrb per launch), maxpoll: polls before a wait
// between workgroups gives up (kPanelWaitTimeout in info), rhs (optional): a
// vector carried through the forward solve (leaves as L^-1 rhs on the panel's rows)
void potrf_panel_f64(hipStream_t s, double* A, int64_t ld, int64_t K0, int nc, int64_t nrb, double* Dinv,
                     double* ldiag, int* info, unsigned* sync, unsigned epoch, unsigned base, unsigned maxpoll,
                     double* rhs);
// the blocked solves of trsv_lower_f64 (N / 64 <= kSolveMaxBlocks) in one launch
// (b is left unchanged); sync: kSolveSyncWords words of the context (zero at
// creation), epoch, base (N / 64 tickets per launch) and maxpoll as for the panel
constexpr int kSolveMaxBlocks = 4096;
constexpr int kSolveSyncWords = kSolveMaxBlocks + 1;
void trsv_persist_f64(hipStream_t s, const double* L, int64_t ld, const double* Dinv, int64_t N, const double* b,
                      double* x, int trans, unsigned* sync, unsigned epoch, unsigned base, int* info,
                      unsigned maxpoll);
// the same on FP32 storage with the FP64 diagonal tiles Dacc (mixed precision)
void potrf_panel_mixed(hipStream_t s, float* A, int64_t ld, int64_t K0, int nc, int64_t nrb, double* Dinv,
                       double* ldiag, int* info, unsigned* sync, unsigned epoch, unsigned base, double* Dacc,
                       unsigned maxpoll);
// X (N x N, ld N), columns c0 .. c0 + ncols: zero above the 64x64 diagonal
// blocks, the blocks from Dinv
void trtri_leaves_f64(hipStream_t s, double* X, int64_t N, const double* Dinv, int64_t c0, int64_t ncols);
void trtri_leaves_f32_batched(hipStream_t s, float* X, int64_t N, const double* Dinv, int batch, int64_t sX,
                              int64_t sD, int64_t c0, int64_t ncols);
// the top level of the recursive-doubling inverse: the largest 64 * 2^k below N (0 for N = 64)
inline int64_t doubling_top(int64_t N) {
    int64_t b = 0;
    for (int64_t t = 64; t < N; t *= 2) b = t;
    return b;
}
void fill_f64(hipStream_t s, double* p, int64_t n, double v);
void scale_rows_f64(hipStream_t s, double* A, int64_t ld, int64_t nrow, int64_t ncol, const double* sv);
// gp_load: saved C_chol (in the n x n part of L, ld N) -> device factor and sv = 1/sqrt(V)
void restore_factor_f64(hipStream_t s, double* L, int64_t n, int64_t N, const double* V, double* sv);
// inverses and log pivots of the T diagonal 64x64 blocks of a lower-triangular factor
void trtri_diag64_f64(hipStream_t s, const double* L, int64_t ld, int64_t T, double* Dinv, double* ldiag, int* info);
// B (lower tiles) = I + s_i K_ij s_j for i,j < n; identity on padding
void form_B_f64(hipStream_t s, const double* K, double* B, int64_t ld, int64_t n, int64_t N,
                const double* sv);
// B64 = I + S K S, B32 = fl32(B64) and the FP64 diagonal tiles (mixed-precision form)
void form_B_mixed(hipStream_t s, const double* K, double* B64, float* B32, double* dacc, int64_t ld, int64_t n,
                  int64_t N, const double* sv);

// Deterministic reductions (lik.hip)
void reduce_sum(hipStream_t s, const double* v, int64_t n, double* out);          // out[0] = sum v
void reduce_dot(hipStream_t s, const double* a, const double* b, int64_t n, double* out);
void reduce_maxabsdiff(hipStream_t s, const double* a, const double* b, int64_t n, double* out);
void reduce_cols(hipStream_t s, const double* part, int64_t rows, int64_t ld, int64_t ncols,
                 double* out);  // out[c] = sum_r part[r*ld + c]

// Likelihood kernels (lik.hip)
// exact != 0: the exact curvature in place of the reference's forward difference
void lik_pseudodata_la(hipStream_t s, int lik, int link, double sigma, int64_t n, const double* f,
                       const double* y, const double* trials, const double* offset, double* z,
                       double* V, double* sv, int* bad, int exact = 0);
// Laplace gradient vectors at the mode: u = d log p / d f, s2 = 1/2 V_i (1 - B^-1_ii) d3_i
void lik_la_grad_vectors(hipStream_t s, int lik, int link, double par, int64_t n, int64_t N, const double* f,
                         const double* y, const double* trials, const double* offset, const double* V,
                         const double* Binv, int64_t ldb, double* u, double* s2);
void lik_loglik(hipStream_t s, int lik, int link, double sigma, int64_t n, const double* f,
                const double* y, const double* trials, const double* offset, double* ll);
// par: the likelihood's parameter (Gaussian sigma, beta-binomial phi)
void ep_sites(hipStream_t s, int lik, int link, double par, int64_t n, const double* fmean, const double* fvar,
              const double* z_old, const double* P_old, const double* y, const double* trials,
              const double* offset, int q, const double* gh_x, const double* gh_w, double damping,
              double* z, double* V, double* sv, double* logZs, double* cav_m, double* cav_v,
              int* nan_flag);
void ep_post(hipStream_t s, int64_t n, const double* V, const double* z, const double* binv_diag,
             const double* cav_m, const double* cav_v, const double* logZs, double* fvar,
             double* term);
void vec_axpby(hipStream_t s, int64_t n, double a, const double* x, double b, const double* y,
               double* out);
void vec_mul(hipStream_t s, int64_t n, const double* a, const double* b, double* out);
void vec_recip(hipStream_t s, int64_t n, const double* a, double* out);

// FITC (fitc.hip)
void fitc_site(hipStream_t s, int64_t n, int64_t N, const double* D, const double* V, const double* z, double* S,
               double* rs, double* zs);
void scale_rows_copy_f64(hipStream_t s, const double* U, int64_t ld, int64_t nrow, int64_t ncol, const double* r,
                         double* W);
void fitc_alpha(hipStream_t s, int64_t n, int64_t N, const double* zs, const double* u, const double* S,
                double* alpha);
void coldot_f64(hipStream_t s, const double* A, const double* B, int64_t ld, int64_t nrow, int64_t ncol, double* out);
void fitc_diag(hipStream_t s, int64_t n, int64_t N, const double* kd, const double* cs, double jitter, double* D);
void vec_log(hipStream_t s, int64_t n, const double* a, double* out);
// C_ii += v_i + jitter (i < n), padded diagonal set to 1
void add_diag_f64(hipStream_t s, double* C, int64_t ld, int64_t n, int64_t N, const double* v, double jitter);
void fitc_ep_var(hipStream_t s, int64_t n, const double* V, const double* S, const double* D, const double* q,
                 double* fvar);
void fitc_train_var(hipStream_t s, int64_t n, const double* prior, const double* t1, const double* Dt,
                    const double* q, const double* S, double* out);
// B = A^T (dimensions multiples of 64)
void transpose_f64(hipStream_t s, const double* A, int64_t lda, int64_t nrow, int64_t ncol, double* B, int64_t ldb);

// Non-log-concave branch (split_kernels.hip)
// out = [rs] K(ridx, cidx) [cs] (+ I) over nr x nc, identity (diag_pad) or zero outside, K
// symmetric in the lower tiles of ldk
void gather_sym_f64(hipStream_t s, const double* K, int64_t ldk, const int64_t* ridx, int64_t nr, const int64_t* cidx,
                    int64_t nc, const double* rs, const double* cs, int add_identity, int diag_pad, double* out,
                    int64_t ld, int64_t NR, int64_t NC);
void gather_vec_f64(hipStream_t s, const double* in, const int64_t* idx, int64_t k, int64_t Kp, double* out);
void scatter_vec_f64(hipStream_t s, const double* in, const int64_t* idx, int64_t k, double* out);
void gather_rows_f64(hipStream_t s, const double* src, int64_t lds, const int64_t* idx, int64_t k, int64_t Kp,
                     int64_t ncol, const double* scale, double* dst, int64_t ldd);
void tril_f64(hipStream_t s, double* A, int64_t ld, int64_t N);
void recip_n_f64(hipStream_t s, const double* a, int64_t k, int64_t Kp, double* out);
void rsqrt_n_f64(hipStream_t s, const double* a, int64_t k, int64_t Kp, double* out);
void ep_split_var(hipStream_t s, int64_t n, const double* V, const double* dinv, double* fvar);
void ep_split_terms(hipStream_t s, int64_t n, int ratio, const double* V, const double* z, const double* cm,
                    const double* cv, const double* logZs, double* term);

// ---- batched forms (one problem per grid row / workgroup, stride per problem) ----
void launch_features_batched(hipStream_t s, const double* x, const int64_t* xoff, const int64_t* nvec,
                             const DevProg* progs, double* F, int64_t N, int64_t sF, int batch);
// first-order log|B| correction for factors of fl32(B) (per lower tile partial sums)
void launch_round_corr(hipStream_t s, const double* B64, const float* W, int64_t N, double* part, int batch,
                       int64_t sB, int64_t sW);
void launch_build_B_f32_batched(hipStream_t s, const DevProg* progs, int nfeat, const double* F, int64_t sF,
                                const int64_t* nvec, int64_t N, double jitter, const double* scal, float* out,
                                int64_t sO, int batch, double* dacc, int64_t sDa, double* out64);
void launch_kmatvec_batched(hipStream_t s, const DevProg* progs, int nfeat, const double* F, int64_t sF,
                            const int64_t* nvec, int64_t N, double jitter, const double* v, double* y, int64_t sV,
                            double* part, int nchunk, int batch);
void launch_grad_tiles_f32_batched(hipStream_t s, const DevProg* progs, const DevProg& hp, const double* F,
                                   int64_t sF, const int64_t* nvec, int64_t N, const float* W, int64_t sW,
                                   const double* wsc, const double* alpha, int64_t sV, double* part, int batch);
void gemm_f32_batched(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, float alpha, const float* A,
                      int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int tiles, int bm,
                      int batch, int64_t sA, int64_t sB, int64_t sC, int fold = 1);
// two batch levels: grid row y moves A, B, C by sA, sB, sC, grid z (nz) all three by sZ
void gemm_f32_batched2(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, float alpha, const float* A,
                       int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int tiles, int bm,
                       int batch, int64_t sA, int64_t sB, int64_t sC, int nz, int64_t sZ, int fold);
// FP64 folding in the GEMMs of L^-1 and B^-1 = X^T X (they feed the gradient and
// the predictive variance, not the log evidence)
constexpr int kFoldInverse = 1;
void gemm_mixed_f32f64_batched(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                               const float* A, int64_t lda, const double* B, int64_t ldb, double beta, float* C,
                               int64_t ldc, int tiles, int batch, int64_t sA, int64_t sB, int64_t sC);
void gemm_mixed_f64f32_batched(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                               const double* A, int64_t lda, const float* B, int64_t ldb, double beta, float* C,
                               int64_t ldc, int tiles, int batch, int64_t sA, int64_t sB, int64_t sC);
void potrf_diag3_f32_batched(hipStream_t s, float* A, int64_t ld, int64_t k, double* Dinv, double* ldiag, int* info,
                             int batch, int64_t sA, int64_t sD, int64_t sL, const double* Dacc);
void set_identity_f32_batched(hipStream_t s, float* A, int64_t ld, int64_t N, int batch, int64_t sA);
void trmv_lower_f32_batched(hipStream_t s, const float* X, int64_t ld, int64_t N, const double* b, double* y,
                            int trans, int batch, int64_t sX, int64_t sv, double* part, int nchunk);
void colsumsq_f32_batched(hipStream_t s, const float* A, int64_t ld, int64_t nrow, int64_t ncol, double* out,
                          int lower, int batch, int64_t sA, int64_t sO);
void seg_scale(hipStream_t s, const double* in, const double* scal, const int64_t* nvec, int64_t N, double* out,
               int batch);
void seg_dot(hipStream_t s, const double* a, const double* b, const int64_t* nvec, int64_t N, double* out,
             int ostride, int slot, int batch);
void seg_cols(hipStream_t s, const double* part, int64_t ntiles, int ng, double* out, int ostride, int slot0,
              int batch);
// per problem: last refinement correction relative to the solution (NaN if non-finite)
void seg_relmax(hipStream_t s, const double* d, const double* u, const int64_t* nvec, int64_t N, double* out,
                int batch);
void batch_finish(hipStream_t s, const double* red, int rs, const int64_t* nvec, const double* sigma,
                  const int32_t* idx, int ng, int nb, double* logz, double* grad, int gstride);

}  // namespace gpl
