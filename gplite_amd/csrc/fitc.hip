// Elementwise and column-reduction kernels of the FITC path (fitc_host.cpp): the
// per-site combination S = D + V of the inversion lemma, the row-scaled copy
// of K_xz, the alpha assembly, the FITC diagonal, and column dot products for
// the predictive variance.  Everything else in FITC runs on the dense kernels.
#include "gpl_internal.h"

#include <cmath>

namespace gpl {

// i < n: S = D + V, rs = 1/sqrt(S), zs = z / S; padding (n <= i < N): zeros
__global__ void k_fitc_site(int64_t n, int64_t N, const double* __restrict__ D, const double* __restrict__ V,
                            const double* __restrict__ z, double* __restrict__ S, double* __restrict__ rs,
                            double* __restrict__ zs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (i < n) {
        const double s = D[i] + V[i];
        S[i] = s;
        rs[i] = 1.0 / sqrt(s);
        zs[i] = z[i] / s;
    } else {
        S[i] = 0.0;
        rs[i] = 0.0;
        zs[i] = 0.0;
    }
}

void fitc_site(hipStream_t s, int64_t n, int64_t N, const double* D, const double* V, const double* z, double* S,
               double* rs, double* zs) {
    hipLaunchKernelGGL(k_fitc_site, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, n, N, D, V, z, S, rs, zs);
}

// W[i + j*ld] = r[i] * U[i + j*ld] over nrow x ncol (one workgroup per column)
__global__ __launch_bounds__(256) void k_scale_rows_copy(const double* __restrict__ U, int64_t ld, int64_t nrow,
                                                         const double* __restrict__ r, double* __restrict__ W) {
    const int64_t j = blockIdx.x;
    for (int64_t i = threadIdx.x; i < nrow; i += 256) W[i + j * ld] = r[i] * U[i + j * ld];
}

void scale_rows_copy_f64(hipStream_t s, const double* U, int64_t ld, int64_t nrow, int64_t ncol, const double* r,
                         double* W) {
    hipLaunchKernelGGL(k_scale_rows_copy, dim3((unsigned)ncol), dim3(256), 0, s, U, ld, nrow, r, W);
}

// i < n: alpha = zs - u / S (R/linear_algebra.R inv_lemma_solve, vector rhs);
// padding: 0
__global__ void k_fitc_alpha(int64_t n, int64_t N, const double* __restrict__ zs, const double* __restrict__ u,
                             const double* __restrict__ S, double* __restrict__ alpha) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    alpha[i] = (i < n) ? zs[i] - u[i] / S[i] : 0.0;
}

void fitc_alpha(hipStream_t s, int64_t n, int64_t N, const double* zs, const double* u, const double* S,
                double* alpha) {
    hipLaunchKernelGGL(k_fitc_alpha, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, n, N, zs, u, S, alpha);
}

// out[j] = sum_i A[i + j*ld] B[i + j*ld], i < nrow (one wavefront per column, fixed order)
__global__ __launch_bounds__(64) void k_coldot(const double* __restrict__ A, const double* __restrict__ B, int64_t ld,
                                               int64_t nrow, double* __restrict__ out) {
    const int64_t j = blockIdx.x;
    const int l = threadIdx.x;
    double acc = 0.0;
    for (int64_t i = l; i < nrow; i += 64) acc += A[i + j * ld] * B[i + j * ld];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (l == 0) out[j] = acc;
}

void coldot_f64(hipStream_t s, const double* A, const double* B, int64_t ld, int64_t nrow, int64_t ncol, double* out) {
    if (ncol <= 0) return;
    hipLaunchKernelGGL(k_coldot, dim3((unsigned)ncol), dim3(64), 0, s, A, B, ld, nrow, out);
}

// FITC diagonal (R/gp_fit.R:107-108): D = (k(x_i, x_i) + jitter) - cs_i for
// i < n, 0 on padding
__global__ void k_fitc_diag(int64_t n, int64_t N, const double* __restrict__ kd, const double* __restrict__ cs,
                            double jitter, double* __restrict__ D) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    D[i] = (i < n) ? (kd[i] + jitter) - cs[i] : 0.0;
}

void fitc_diag(hipStream_t s, int64_t n, int64_t N, const double* kd, const double* cs, double jitter, double* D) {
    hipLaunchKernelGGL(k_fitc_diag, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, n, N, kd, cs, jitter, D);
}

// out[i] = log a[i], i < n
__global__ void k_vec_log(int64_t n, const double* __restrict__ a, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = log(a[i]);
}

void vec_log(hipStream_t s, int64_t n, const double* a, double* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_vec_log, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, out);
}

// B (ncol x nrow, ld ldb) = A^T for A (nrow x ncol, ld lda); both dimensions
// multiples of 64.  64x64 tiles through LDS so reads and writes are coalesced.
__global__ __launch_bounds__(256) void k_transpose64(const double* __restrict__ A, int64_t lda, int64_t tiles_r,
                                                     double* __restrict__ B, int64_t ldb) {
    __shared__ double t[64][65];
    const int64_t ti = blockIdx.x % tiles_r, tj = blockIdx.x / tiles_r;
    const int e = threadIdx.x & 63, p = threadIdx.x >> 6;
    const double* a = A + ti * 64 + tj * 64 * lda;
    for (int c = p; c < 64; c += 4) t[c][e] = a[e + (int64_t)c * lda];  // t[col][row]
    __syncthreads();
    double* b = B + tj * 64 + ti * 64 * ldb;
    for (int c = p; c < 64; c += 4) b[e + (int64_t)c * ldb] = t[e][c];  // B(tj*64+e, ti*64+c) = A(ti*64+c, tj*64+e)
}

void transpose_f64(hipStream_t s, const double* A, int64_t lda, int64_t nrow, int64_t ncol, double* B, int64_t ldb) {
    const int64_t tr = nrow / 64, tc = ncol / 64;
    if (tr <= 0 || tc <= 0) return;
    hipLaunchKernelGGL(k_transpose64, dim3((unsigned)(tr * tc)), dim3(256), 0, s, A, lda, tr, B, ldb);
}

// C[i + i*ld] += v[i] + jitter for i < n; the padded diagonal (n <= i < N) becomes 1
__global__ void k_add_diag(double* __restrict__ C, int64_t ld, int64_t n, int64_t N, const double* __restrict__ v,
                           double jitter) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (i < n) C[i + i * ld] += v[i] + jitter;
    else C[i + i * ld] = 1.0;
}

void add_diag_f64(hipStream_t s, double* C, int64_t ld, int64_t n, int64_t N, const double* v, double jitter) {
    hipLaunchKernelGGL(k_add_diag, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, C, ld, n, N, v, jitter);
}

// EP posterior variance of FITC (R/ep.R:170-174 through the inversion lemma):
// fvar = V [K_f C^-1]_ii = V (q/S + D (1/S - q/S^2)), q_i = |Lc^-1 u_i|^2
__global__ void k_fitc_ep_var(int64_t n, const double* __restrict__ V, const double* __restrict__ S,
                              const double* __restrict__ D, const double* __restrict__ q, double* __restrict__ fvar) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double si = S[i], qi = q[i];
    fvar[i] = V[i] * (qi / si + D[i] * (1.0 / si - qi / (si * si)));
}

void fitc_ep_var(hipStream_t s, int64_t n, const double* V, const double* S, const double* D, const double* q,
                 double* fvar) {
    hipLaunchKernelGGL(k_fitc_ep_var, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, V, S, D, q, fvar);
}

// Training-input variance (R/gp_pred.R:270-285): prior - (t1 + 2 Dt q/S + Dt^2 (1/S - q/S^2)),
// t1 = (A^T M A)_ii, the three terms diag1 + diag2 + diag3 of the reference
__global__ void k_fitc_train_var(int64_t n, const double* __restrict__ prior, const double* __restrict__ t1,
                                 const double* __restrict__ Dt, const double* __restrict__ q,
                                 const double* __restrict__ S, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double si = S[i], qi = q[i], d = Dt[i];
    out[i] = prior[i] - (t1[i] + 2.0 * d * qi / si + d * d * (1.0 / si - qi / (si * si)));
}

void fitc_train_var(hipStream_t s, int64_t n, const double* prior, const double* t1, const double* Dt,
                    const double* q, const double* S, double* out) {
    hipLaunchKernelGGL(k_fitc_train_var, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, prior, t1, Dt, q, S,
                       out);
}

}  // namespace gpl
