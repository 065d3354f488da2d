// Gathers and small elementwise kernels for the non-log-concave branch: the
// sites are split into those with positive pseudo-variance ("good") and the
// rest ("bad"), and the blocks of K that the branch needs are gathered from the
// symmetric K held in its lower 64x64 tiles (K[a + b*ld] for a >= b).
#include "gpl_internal.h"

#include <algorithm>

namespace gpl {

__device__ __forceinline__ double sym_at(const double* __restrict__ K, int64_t ld, int64_t a, int64_t b) {
    return a >= b ? K[a + b * ld] : K[b + a * ld];
}

// out[i + j*ld] = rs_i K(r_i, c_j) cs_j (+1 on i == j when add_identity), for
// i < nr and j < nc; outside that block out is the identity (diag_pad) or zero.
// rs / cs may be null (no scaling).
__global__ void k_gather_sym(const double* __restrict__ K, int64_t ldk, const int64_t* __restrict__ ridx, int64_t nr,
                             const int64_t* __restrict__ cidx, int64_t nc, const double* __restrict__ rs,
                             const double* __restrict__ cs, int add_identity, int diag_pad, double* __restrict__ out,
                             int64_t ld, int64_t NR, int64_t NC) {
    const int64_t i = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int64_t j = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= NR || j >= NC) return;
    double v;
    if (i < nr && j < nc) {
        v = sym_at(K, ldk, ridx[i], cidx[j]);
        if (rs) v *= rs[i];
        if (cs) v *= cs[j];
        if (add_identity && i == j) v += 1.0;
    } else {
        v = (diag_pad && i == j) ? 1.0 : 0.0;
    }
    out[i + j * ld] = v;
}

void gather_sym_f64(hipStream_t s, const double* K, int64_t ldk, const int64_t* ridx, int64_t nr, const int64_t* cidx,
                    int64_t nc, const double* rs, const double* cs, int add_identity, int diag_pad, double* out,
                    int64_t ld, int64_t NR, int64_t NC) {
    if (NR <= 0 || NC <= 0) return;
    hipLaunchKernelGGL(k_gather_sym, dim3((unsigned)((NR + 63) / 64), (unsigned)((NC + 3) / 4)), dim3(256), 0, s, K,
                       ldk, ridx, nr, cidx, nc, rs, cs, add_identity, diag_pad, out, ld, NR, NC);
}

// out[i] = scale_i * in[idx[i]] for i < k, 0 for k <= i < K (scale may be null)
__global__ void k_gather_vec(const double* __restrict__ in, const int64_t* __restrict__ idx, int64_t k, int64_t Kp,
                             double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Kp) return;
    out[i] = i < k ? in[idx[i]] : 0.0;
}

void gather_vec_f64(hipStream_t s, const double* in, const int64_t* idx, int64_t k, int64_t Kp, double* out) {
    if (Kp <= 0) return;
    hipLaunchKernelGGL(k_gather_vec, dim3((unsigned)((Kp + 255) / 256)), dim3(256), 0, s, in, idx, k, Kp, out);
}

// out[idx[i]] = in[i] for i < k
__global__ void k_scatter_vec(const double* __restrict__ in, const int64_t* __restrict__ idx, int64_t k,
                              double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) out[idx[i]] = in[i];
}

void scatter_vec_f64(hipStream_t s, const double* in, const int64_t* idx, int64_t k, double* out) {
    if (k <= 0) return;
    hipLaunchKernelGGL(k_scatter_vec, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, s, in, idx, k, out);
}

// rows idx[0..k) of src (leading dimension lds, ncol columns) into rows 0..k of dst
// (leading dimension ldd), rows k..Kp zero; scaled by scale_i when given
__global__ void k_gather_rows(const double* __restrict__ src, int64_t lds, const int64_t* __restrict__ idx, int64_t k,
                              int64_t Kp, const double* __restrict__ scale, double* __restrict__ dst, int64_t ldd) {
    const int64_t j = blockIdx.y;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < Kp; i += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        if (i < k) {
            v = src[idx[i] + j * lds];
            if (scale) v *= scale[i];
        }
        dst[i + j * ldd] = v;
    }
}

void gather_rows_f64(hipStream_t s, const double* src, int64_t lds, const int64_t* idx, int64_t k, int64_t Kp,
                     int64_t ncol, const double* scale, double* dst, int64_t ldd) {
    if (Kp <= 0 || ncol <= 0) return;
    const unsigned gx = (unsigned)std::min<int64_t>((Kp + 255) / 256, 64);
    hipLaunchKernelGGL(k_gather_rows, dim3(gx, (unsigned)ncol), dim3(256), 0, s, src, lds, idx, k, Kp, scale, dst, ldd);
}

// zero every entry above the diagonal of an N x N matrix (a factor written in
// the lower tiles keeps B's values above the diagonal of its diagonal tiles)
__global__ void k_tril(double* __restrict__ A, int64_t ld, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int64_t j = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i < N && j < N && j > i) A[i + j * ld] = 0.0;
}

void tril_f64(hipStream_t s, double* A, int64_t ld, int64_t N) {
    if (N <= 0) return;
    hipLaunchKernelGGL(k_tril, dim3((unsigned)((N + 63) / 64), (unsigned)((N + 3) / 4)), dim3(256), 0, s, A, ld, N);
}

// out[i] = 1 / a[i] for i < k, 0 for k <= i < Kp
__global__ void k_recip_n(const double* __restrict__ a, int64_t k, int64_t Kp, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Kp) out[i] = i < k ? 1.0 / a[i] : 0.0;
}

void recip_n_f64(hipStream_t s, const double* a, int64_t k, int64_t Kp, double* out) {
    if (Kp <= 0) return;
    hipLaunchKernelGGL(k_recip_n, dim3((unsigned)((Kp + 255) / 256)), dim3(256), 0, s, a, k, Kp, out);
}

// 1 / sqrt(v[i]) for i < k, 0 for k <= i < Kp
__global__ void k_rsqrt_n(const double* __restrict__ a, int64_t k, int64_t Kp, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Kp) out[i] = i < k ? sqrt(1.0 / a[i]) : 0.0;
}

void rsqrt_n_f64(hipStream_t s, const double* a, int64_t k, int64_t Kp, double* out) {
    if (Kp <= 0) return;
    hipLaunchKernelGGL(k_rsqrt_n, dim3((unsigned)((Kp + 255) / 256)), dim3(256), 0, s, a, k, Kp, out);
}

// EP, split sites: fvar_i = V_i (1 - V_i [(K + V)^-1]_ii) (the diagonal of
// diag(V) (K + V)^-1 K, since (K + V)^-1 K = I - (K + V)^-1 V)
__global__ void k_ep_split_var(int64_t n, const double* __restrict__ V, const double* __restrict__ dinv,
                               double* __restrict__ fvar) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) fvar[i] = V[i] * (1.0 - V[i] * dinv[i]);
}

void ep_split_var(hipStream_t s, int64_t n, const double* V, const double* dinv, double* fvar) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ep_split_var, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, V, dinv, fvar);
}

// EP evidence terms per site (R/ep.R:45-48 for the sites with V > 0, R/ep.R:104-107
// for the others): logZ_i + 1/2 log(c_i + V_i) - 1/2 log V_i + 1/2 (m_i - z_i)^2 / (c_i + V_i),
// with the first two logs as one log((c_i + V_i) / V_i) when `ratio` is set (V_i <= 0)
__global__ void k_ep_split_terms(int64_t n, int ratio, const double* __restrict__ V, const double* __restrict__ z,
                                 const double* __restrict__ cm, const double* __restrict__ cv,
                                 const double* __restrict__ logZs, double* __restrict__ term) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double Vi = V[i], c = cv[i] + Vi, dz = cm[i] - z[i];
    const double lg = ratio ? 0.5 * log(c / Vi) : 0.5 * log(c) - 0.5 * log(Vi);
    term[i] = logZs[i] + lg + 0.5 * dz * dz / c;
}

void ep_split_terms(hipStream_t s, int64_t n, int ratio, const double* V, const double* z, const double* cm,
                    const double* cv, const double* logZs, double* term) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ep_split_terms, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, ratio, V, z, cm, cv,
                       logZs, term);
}

}  // namespace gpl
