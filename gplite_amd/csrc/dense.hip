// Dense FP64 linear algebra for gfx950: MFMA GEMM, 64x64 diagonal-block Cholesky
// with in-LDS inverse, blocked triangular solves and matrix-vector products.
//
// Every matrix is column-major with a leading dimension that is a multiple of 64
// and logical sizes padded to multiples of 64 (padding is identity / zero), so the
// kernels only ever see whole 64x64 tiles.
#include "gpl_internal.h"

#include <cmath>

namespace gpl {

typedef double d4 __attribute__((ext_vector_type(4)));

// Map the hardware block id so that each XCD (blocks are dealt round-robin over
// 8 XCDs) receives a contiguous range of logical tiles: neighbouring tiles share
// operand panels, which then stay in that XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nb) {
    const int64_t per = nb / 8, rem = nb % 8;
    const int64_t x = b % 8, q = b / 8;
    return (x < rem) ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}

__device__ __forceinline__ void lower_tile(int64_t t, int64_t& ti, int64_t& tj) {
    int64_t r = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    ti = r;
    tj = t - r * (r + 1) / 2;
}

// ---------------------------------------------------------------------------
// GEMM: C = beta*C + alpha * op(A) op(B).  64x64 output tile per 256-thread
// workgroup; each wave owns a 32x32 quadrant as 2x2 v_mfma_f64_16x16x4 blocks.
// Operands are read straight from global memory in MFMA operand layout.  Within
// a 16-wide k chunk, lane group g = lane/16 takes k = 4g..4g+3 over the four
// MFMA steps, so a transposed operand is a contiguous 32-byte read per lane and
// a non-transposed one a 128-byte coalesced read per 16 lanes.
// ---------------------------------------------------------------------------
template <int TA, int TB>
__device__ __forceinline__ void gemm_load(const double* A, int64_t lda,
                                          const double* B, int64_t ldb, int64_t row0,
                                          int64_t col0, int64_t k0, int li, int g, double (&a)[2][4],
                                          double (&b)[2][4]) {
    const int64_t kk = k0 + 4 * g;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        const int64_t r = row0 + bi * 16 + li;
        if (TA) {
            const double2* p = reinterpret_cast<const double2*>(A + kk + r * lda);
            double2 u = p[0], v = p[1];
            a[bi][0] = u.x; a[bi][1] = u.y; a[bi][2] = v.x; a[bi][3] = v.y;
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) a[bi][s] = A[r + (kk + s) * lda];
        }
    }
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
        const int64_t c = col0 + bj * 16 + li;
        if (!TB) {
            const double2* p = reinterpret_cast<const double2*>(B + kk + c * ldb);
            double2 u = p[0], v = p[1];
            b[bj][0] = u.x; b[bj][1] = u.y; b[bj][2] = v.x; b[bj][3] = v.y;
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) b[bj][s] = B[c + (kk + s) * ldb];
        }
    }
}

__device__ __forceinline__ void gemm_mma(const double (&a)[2][4], const double (&b)[2][4], d4 (&acc)[2][2]) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
                acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi][s], b[bj][s], acc[bi][bj], 0, 0, 0);
}

template <int TA, int TB>
__global__ __launch_bounds__(256) void k_gemm_f64(int64_t tilesM, int64_t K, double alpha,
                                                  const double* A, int64_t lda,
                                                  const double* B, int64_t ldb, double beta,
                                                  double* C, int64_t ldc, int tiles) {
    const int64_t nb = gridDim.x;
    const int64_t bid = xcd_remap(blockIdx.x, nb);
    int64_t ti, tj;
    if (tiles == 0 || tiles == 3) {
        ti = bid % tilesM;
        tj = bid / tilesM;
    } else {
        lower_tile(bid, ti, tj);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int64_t row0 = ti * 64 + (wave >> 1) * 32;
    const int64_t col0 = tj * 64 + (wave & 1) * 32;
    const int64_t kbeg = (tiles == 2) ? ti * 64 : 0;
    if (tiles == 3 && (ti + 1) * 64 < K) K = (ti + 1) * 64;

    d4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    double a0[2][4], b0[2][4], a1[2][4], b1[2][4];
    int64_t k0 = kbeg;
    if (k0 < K) gemm_load<TA, TB>(A, lda, B, ldb, row0, col0, k0, li, g, a0, b0);
    while (k0 < K) {
        if (k0 + 16 < K) gemm_load<TA, TB>(A, lda, B, ldb, row0, col0, k0 + 16, li, g, a1, b1);
        gemm_mma(a0, b0, acc);
        k0 += 16;
        if (k0 >= K) break;
        if (k0 + 16 < K) gemm_load<TA, TB>(A, lda, B, ldb, row0, col0, k0 + 16, li, g, a0, b0);
        gemm_mma(a1, b1, acc);
        k0 += 16;
    }

    // C may alias an operand (in-place panel solve): every wave finishes reading first.
    __syncthreads();
    // v_mfma_f64_16x16x4 D layout: lane holds rows g, g+4, g+8, g+12 of column li
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const int64_t c = col0 + bj * 16 + li;
            const int64_t r = row0 + bi * 16 + g;
            double* p = C + r + c * ldc;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double o = alpha * acc[bi][bj][v];
                p[4 * v] = (beta == 0.0) ? o : beta * p[4 * v] + o;
            }
        }
}

void gemm_f64(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
              const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
              int64_t ldc, int tiles) {
    const int64_t tm = M / 64, tn = N / 64;
    if (tm <= 0 || tn <= 0) return;
    const int64_t nblk = (tiles == 0 || tiles == 3) ? tm * tn : tm * (tm + 1) / 2;
    dim3 grid((unsigned)nblk), block(256);
    if (K <= 0) {  // pure scaling (alpha * 0 + beta*C)
        if (beta == 1.0) return;
    }
    if (!ta && !tb)
        hipLaunchKernelGGL((k_gemm_f64<0, 0>), grid, block, 0, s, tm, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles);
    else if (!ta && tb)
        hipLaunchKernelGGL((k_gemm_f64<0, 1>), grid, block, 0, s, tm, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles);
    else if (ta && !tb)
        hipLaunchKernelGGL((k_gemm_f64<1, 0>), grid, block, 0, s, tm, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles);
    else
        hipLaunchKernelGGL((k_gemm_f64<1, 1>), grid, block, 0, s, tm, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles);
}

// ---------------------------------------------------------------------------
// Diagonal block: right-looking Cholesky of a 64x64 block held in LDS (one
// barrier per column), then its inverse by column-parallel substitution with
// four lanes per column exchanging pivots through cross-lane shuffles.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_potrf_diag(double* __restrict__ A, int64_t ld, int64_t k,
                                                    double* __restrict__ Dinv, double* __restrict__ ldiag,
                                                    int* __restrict__ info) {
    __shared__ double a[64][65];  // a[col][row]
    __shared__ double rinv[64];
    double* blk = A + k * 64 * (ld + 1);
    const int t = threadIdx.x;
    const int i = t & 63, grp = t >> 6;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = grp + 4 * q;
        a[c][i] = (i >= c) ? blk[i + (int64_t)c * ld] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < 64; ++j) {
        const double d = a[j][j];
        const double r = sqrt(d);
        const double inv = 1.0 / r;
        if (t == 0 && !(d > 0.0)) atomicCAS(info, 0, (int)(k * 64 + j + 1));
        const double lij = a[j][i] * inv;
        if (i > j) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int c = grp + 4 * q;
                if (c > j && c <= i) a[c][i] -= lij * (a[j][c] * inv);
            }
        }
        __syncthreads();
        if (grp == 0) {
            if (i > j) a[j][i] = lij;
            else if (i == j) { a[j][j] = r; rinv[j] = inv; }
        }
    }
    __syncthreads();
    // store L (upper part of the tile zeroed) and log of the diagonal
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = grp + 4 * q;
        blk[i + (int64_t)c * ld] = (i >= c) ? a[c][i] : 0.0;
    }
    if (t < 64) ldiag[k * 64 + t] = log(a[t][t]);
    __syncthreads();
    if (t < 64) a[t][t] = 0.0;  // strictly lower part drives the substitution below
    __syncthreads();

    // inverse: column c = 16*wave + lane/4; this lane owns rows 4r+q
    const int lane = t & 63;
    const int c = grp * 16 + (lane >> 2);
    const int q = lane & 3;
    double b[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) b[r] = (4 * r + q == c) ? 1.0 : 0.0;
#pragma unroll
    for (int kk = 0; kk < 64; ++kk) {
        const int kr = kk >> 2, kq = kk & 3;
        double xk = (q == kq) ? b[kr] * rinv[kk] : 0.0;
        xk = __shfl(xk, (lane & ~3) | kq, 64);
        if (q == kq) b[kr] = xk;
#pragma unroll
        for (int r = 0; r < 16; ++r) b[r] -= a[kk][4 * r + q] * xk;
    }
    double* D = Dinv + k * 64 * 64;
#pragma unroll
    for (int r = 0; r < 16; ++r) D[(4 * r + q) + c * 64] = b[r];
}

void potrf_diag_f64(hipStream_t s, double* A, int64_t ld, int64_t k, double* Dinv, double* ldiag, int* info) {
    hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(256), 0, s, A, ld, k, Dinv, ldiag, info);
}

// ---------------------------------------------------------------------------
// Blocked triangular solves with one vector (L y = b, L^T x = b).  Step i:
// every workgroup forms x_i = Dinv_i b_i (Dinv_i^T b_i for the transpose) and
// workgroup w updates the pending block r of b with the tile L(r,i) (or L(i,r)^T).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum4(double v, double* red) {
    // 256 threads: thread t holds the partial of element (t & 63), part (t >> 6)
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    double s = 0.0;
    if (t < 64) s = red[t] + red[t + 64] + red[t + 128] + red[t + 192];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(256) void k_trsv_step(const double* __restrict__ L, int64_t ld,
                                                   const double* __restrict__ Dinv, int64_t T, int64_t i,
                                                   double* __restrict__ b, double* __restrict__ x, int trans) {
    __shared__ double red[256];
    __shared__ double xi[64];
    const int t = threadIdx.x, e = t & 63, p = t >> 6;
    const double* D = Dinv + i * 64 * 64;
    // x_i = D b_i (trans: D^T b_i); element e, columns/rows p*16..p*16+15
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int c = p * 16 + u;
        s += (trans ? D[c + e * 64] : D[e + c * 64]) * b[i * 64 + c];
    }
    s = block_sum4(s, red);
    if (t < 64) xi[t] = s;
    __syncthreads();
    if (blockIdx.x == 0 && t < 64) x[i * 64 + t] = xi[t];
    // pending block for this workgroup
    const int64_t w = blockIdx.x;
    int64_t r;
    if (!trans) { r = i + 1 + w; if (r >= T) return; }
    else        { r = i - 1 - w; if (r < 0) return; }
    double u2 = 0.0;
    if (!trans) {
        const double* Lt = L + r * 64 + i * 64 * ld;  // tile (r, i)
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int c = p * 16 + u;
            u2 += Lt[e + (int64_t)c * ld] * xi[c];
        }
    } else {
        const double* Lt = L + i * 64 + r * 64 * ld;  // tile (i, r): (L^T)(r,i) = tile^T
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int c = p * 16 + u;
            u2 += Lt[c + (int64_t)e * ld] * xi[c];
        }
    }
    u2 = block_sum4(u2, red);
    if (t < 64) b[r * 64 + t] -= u2;
}

void trsv_lower_f64(hipStream_t s, const double* L, int64_t ld, const double* Dinv, int64_t N, double* b,
                    double* x, int trans) {
    // b is consumed as workspace; the solution is written to x.
    const int64_t T = N / 64;
    for (int64_t st = 0; st < T; ++st) {
        const int64_t i = trans ? (T - 1 - st) : st;
        const int64_t pend = trans ? i : (T - 1 - i);
        hipLaunchKernelGGL(k_trsv_step, dim3((unsigned)(pend > 0 ? pend : 1)), dim3(256), 0, s, L, ld, Dinv,
                           T, i, b, x, trans);
    }
}

// ---------------------------------------------------------------------------
// Triangular / symmetric matrix-vector products over lower tiles.  One
// workgroup per 64-row output block, fixed summation order (deterministic).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_trmv(const double* __restrict__ X, int64_t ld, int64_t T,
                                              const double* __restrict__ b, double* __restrict__ y,
                                              int trans) {
    __shared__ double red[256];
    const int64_t ti = blockIdx.x;
    const int t = threadIdx.x, e = t & 63, p = t >> 6;
    double s = 0.0;
    if (!trans) {
        // y[ti] = sum_{tj <= ti} X(ti,tj) b[tj]
        for (int64_t tj = 0; tj <= ti; ++tj) {
            const double* Xt = X + ti * 64 + tj * 64 * ld;
#pragma unroll 4
            for (int u = 0; u < 16; ++u) {
                const int c = p * 16 + u;
                s += Xt[e + (int64_t)c * ld] * b[tj * 64 + c];
            }
        }
    } else {
        // y[ti] = sum_{tr >= ti} X(tr,ti)^T b[tr]; element e = column of the tile
        for (int64_t tr = ti; tr < T; ++tr) {
            const double* Xt = X + tr * 64 + ti * 64 * ld;
#pragma unroll 4
            for (int u = 0; u < 16; ++u) {
                const int r = p * 16 + u;
                s += Xt[r + (int64_t)e * ld] * b[tr * 64 + r];
            }
        }
    }
    s = block_sum4(s, red);
    if (t < 64) y[ti * 64 + t] = s;
}

void trmv_lower_f64(hipStream_t s, const double* X, int64_t ld, int64_t N, const double* b, double* y, int trans,
                    double* part) {
    (void)part;
    const int64_t T = N / 64;
    hipLaunchKernelGGL(k_trmv, dim3((unsigned)T), dim3(256), 0, s, X, ld, T, b, y, trans);
}

__global__ __launch_bounds__(256) void k_symv(const double* __restrict__ K, int64_t ld, int64_t T,
                                              const double* __restrict__ v, double* __restrict__ y) {
    __shared__ double red[256];
    const int64_t ti = blockIdx.x;
    const int t = threadIdx.x, e = t & 63, p = t >> 6;
    double s = 0.0;
    for (int64_t tj = 0; tj < T; ++tj) {
        if (tj <= ti) {
            const double* Kt = K + ti * 64 + tj * 64 * ld;
#pragma unroll 4
            for (int u = 0; u < 16; ++u) {
                const int c = p * 16 + u;
                const double kv = (tj == ti && c > e) ? Kt[c + (int64_t)e * ld] : Kt[e + (int64_t)c * ld];
                s += kv * v[tj * 64 + c];
            }
        } else {
            const double* Kt = K + tj * 64 + ti * 64 * ld;  // tile (tj, ti), use transposed
#pragma unroll 4
            for (int u = 0; u < 16; ++u) {
                const int c = p * 16 + u;
                s += Kt[c + (int64_t)e * ld] * v[tj * 64 + c];
            }
        }
    }
    s = block_sum4(s, red);
    if (t < 64) y[ti * 64 + t] = s;
}

void symv_lower_f64(hipStream_t s, const double* K, int64_t ld, int64_t N, const double* v, double* y,
                    double* part) {
    (void)part;
    hipLaunchKernelGGL(k_symv, dim3((unsigned)(N / 64)), dim3(256), 0, s, K, ld, N / 64, v, y);
}

// one wave per column: out[j] = sum_i A[i + j*ld]^2; lower != 0 starts at the
// column's diagonal tile (the upper tiles of a lower-triangular buffer are never written)
__global__ __launch_bounds__(256) void k_colsumsq(const double* __restrict__ A, int64_t ld, int64_t nrow,
                                                  int64_t ncol, double* __restrict__ out, int lower) {
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= ncol) return;
    const double* col = A + j * ld;
    double s = 0.0;
    for (int64_t i = (lower ? (j / 64) * 64 : 0) + lane; i < nrow; i += 64) {
        const double v = col[i];
        s += v * v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) out[j] = s;
}

__global__ __launch_bounds__(256) void k_gemv_t(const double* __restrict__ A, int64_t ld, int64_t nrow,
                                                int64_t ncol, const double* __restrict__ v, double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= ncol) return;
    const double* col = A + j * ld;
    double s = 0.0;
    for (int64_t i = lane; i < nrow; i += 64) s += col[i] * v[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) out[j] = s;
}

void gemv_t_f64(hipStream_t s, const double* A, int64_t ld, int64_t nrow, int64_t ncol, const double* v,
                double* out) {
    if (ncol <= 0) return;
    hipLaunchKernelGGL(k_gemv_t, dim3((unsigned)((ncol + 3) / 4)), dim3(256), 0, s, A, ld, nrow, ncol, v, out);
}

void colsumsq_f64(hipStream_t s, const double* A, int64_t ld, int64_t nrow, int64_t ncol, double* out, int lower) {
    if (ncol <= 0) return;
    hipLaunchKernelGGL(k_colsumsq, dim3((unsigned)((ncol + 3) / 4)), dim3(256), 0, s, A, ld, nrow, ncol, out, lower);
}

// ---------------------------------------------------------------------------
// Elementwise helpers over lower tiles / vectors
// ---------------------------------------------------------------------------
__global__ void k_set_identity_lower(double* __restrict__ A, int64_t ld, int64_t T) {
    int64_t ti, tj;
    lower_tile(blockIdx.x, ti, tj);
    const int t = threadIdx.x, i = t & 63;
    for (int c = t >> 6; c < 64; c += 4) {
        const int64_t gi = ti * 64 + i, gj = tj * 64 + c;
        A[gi + gj * ld] = (gi == gj) ? 1.0 : 0.0;
    }
}

void set_identity_lower_f64(hipStream_t s, double* A, int64_t ld, int64_t N) {
    const int64_t T = N / 64;
    hipLaunchKernelGGL(k_set_identity_lower, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, A, ld, T);
}

__global__ void k_fill(double* p, int64_t n, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void fill_f64(hipStream_t s, double* p, int64_t n, double v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

__global__ void k_scale_rows(double* __restrict__ A, int64_t ld, int64_t nrow, int64_t ncol,
                             const double* __restrict__ sv) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i < nrow) A[i + j * ld] *= sv[i];
}

void scale_rows_f64(hipStream_t s, double* A, int64_t ld, int64_t nrow, int64_t ncol, const double* sv) {
    if (nrow <= 0 || ncol <= 0) return;
    hipLaunchKernelGGL(k_scale_rows, dim3((unsigned)((nrow + 255) / 256), (unsigned)ncol), dim3(256), 0, s, A,
                       ld, nrow, ncol, sv);
}

__global__ void k_form_B(const double* __restrict__ K, double* __restrict__ B, int64_t ld, int64_t n,
                         const double* __restrict__ sv) {
    int64_t ti, tj;
    lower_tile(blockIdx.x, ti, tj);
    const int t = threadIdx.x, i = t & 63;
    const int64_t gi = ti * 64 + i;
    const double si = (gi < n) ? sv[gi] : 0.0;
    for (int c = t >> 6; c < 64; c += 4) {
        const int64_t gj = tj * 64 + c;
        double v;
        if (gi < n && gj < n) v = si * K[gi + gj * ld] * sv[gj] + (gi == gj ? 1.0 : 0.0);
        else v = (gi == gj) ? 1.0 : 0.0;
        B[gi + gj * ld] = v;
    }
}

void form_B_f64(hipStream_t s, const double* K, double* B, int64_t ld, int64_t n, int64_t N, const double* sv) {
    const int64_t T = N / 64;
    hipLaunchKernelGGL(k_form_B, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, K, B, ld, n, sv);
}

}  // namespace gpl
