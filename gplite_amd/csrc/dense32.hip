// FP32 dense kernels for the mixed-precision form on gfx950: an LDS-staged
// GEMM on v_mfma_f32_32x32x2f32 with FP32 accumulation (trailing updates,
// triangular-inverse updates, B^-1 = X^T X), and FP32-matrix / FP64-vector
// helpers used by iterative refinement.
#include "gpl_internal.h"

#include <cmath>

namespace gpl {

typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ int64_t xcd_remap32(int64_t b, int64_t nb) {
    const int64_t per = nb / 8, rem = nb % 8;
    const int64_t x = b % 8, q = b / 8;
    return (x < rem) ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}

__device__ __forceinline__ void lower_tile32(int64_t t, int64_t& ti, int64_t& tj) {
    int64_t r = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    ti = r;
    tj = t - r * (r + 1) / 2;
}

// chunk of KC k x BM m, FP32, into registers (NT threads, KC*BM/(4*NT) float4 each)
template <int BM, int MCONT, int NT, int KC>
__device__ __forceinline__ void g32_load(const float* P, int64_t ld, int64_t base, int64_t lim, int64_t k0,
                                         float4 (&r)[KC * BM / (4 * NT)]) {
    constexpr int AL = KC * BM / (4 * NT);
    const int t = threadIdx.x;
    if (MCONT) {  // (m, k) at P[m + k*ld]: float4 along m
        const int m4 = (t % (BM / 4)) * 4, kq = t / (BM / 4), ks = NT / (BM / 4);
        const bool ok = base + m4 < lim;
#pragma unroll
        for (int j = 0; j < AL; ++j) {
            const int64_t k = k0 + kq + ks * j;
            r[j] = ok ? *reinterpret_cast<const float4*>(P + base + m4 + k * ld) : float4{0.f, 0.f, 0.f, 0.f};
        }
    } else {  // (m, k) at P[k + m*ld]: KC/4 lanes read one column's KC k
        constexpr int LPC = KC / 4, CPP = NT / LPC;  // lanes per column, columns per pass
        const int kq = t % LPC, m = t / LPC;
#pragma unroll
        for (int j = 0; j < AL; ++j) {
            const int64_t mm = base + m + CPP * j;
            r[j] = (mm < lim) ? *reinterpret_cast<const float4*>(P + (k0 + 4 * kq) + mm * ld)
                              : float4{0.f, 0.f, 0.f, 0.f};
        }
    }
}

template <int BM, int MCONT, int NT, int KC>
__device__ __forceinline__ void g32_store(float (*S)[BM + 4], const float4 (&r)[KC * BM / (4 * NT)]) {
    constexpr int AL = KC * BM / (4 * NT);
    const int t = threadIdx.x;
    if (MCONT) {
        const int m4 = (t % (BM / 4)) * 4, kq = t / (BM / 4), ks = NT / (BM / 4);
#pragma unroll
        for (int j = 0; j < AL; ++j) *reinterpret_cast<float4*>(&S[kq + ks * j][m4]) = r[j];
    } else {
        constexpr int LPC = KC / 4, CPP = NT / LPC;
        const int kq = t % LPC, m = t / LPC;
#pragma unroll
        for (int j = 0; j < AL; ++j) {
            S[4 * kq + 0][m + CPP * j] = r[j].x;
            S[4 * kq + 1][m + CPP * j] = r[j].y;
            S[4 * kq + 2][m + CPP * j] = r[j].z;
            S[4 * kq + 3][m + CPP * j] = r[j].w;
        }
    }
}

// C = beta*C + alpha*op(A) op(B), FP32 storage and FP32 MFMA products.  Tile
// modes as gemm2_f64.  NW waves in a 2 x NW/2 grid each own a (BM/2) x (2BM/NW)
// block of 32x32 MFMA tiles; lane half h = lane/32 takes k = (KC/2)h + s of a
// KC-deep chunk over KC/2 MFMA steps.  With FOLD > 0 the FP32 accumulators
// cover at most KC*FOLD consecutive k: they are then added into FP64
// accumulators and cleared, and the epilogue forms beta*C + alpha*acc in FP64
// with one rounding to FP32.  A long FP32 running sum loses about sqrt(k) ulps
// of its partial sums; in trailing updates that cancel (A - P P^T with
// |P P^T| >> |result|) that error dominates log|B|.  FOLD == 0 keeps plain
// FP32 accumulation.
template <int TA, int TB, int BM, int FOLD, int NW, int KC>
__global__ __launch_bounds__(64 * NW) void k_gemm32(int64_t M, int64_t N, int64_t K, int64_t tilesM, float alpha,
                                                    const float* A, int64_t lda, const float* B, int64_t ldb,
                                                    float beta, float* C, int64_t ldc, int tiles, int64_t sA,
                                                    int64_t sB, int64_t sC, int64_t sZ) {
    constexpr int NT = 64 * NW, WCn = NW / 2, WTR = BM / 2, WTC = BM / WCn;
    constexpr int NBR = WTR / 32, NBC = WTC / 32, AL = KC * BM / (4 * NT);
    constexpr int FR = FOLD > 0 ? NBR : 1, FC = FOLD > 0 ? NBC : 1;
    static_assert(NBR >= 1 && NBC >= 1 && AL >= 1, "tile shape");
    // batch: one problem per grid row; a second batch level (grid z) moves all
    // three operands by sZ
    A += blockIdx.y * sA + blockIdx.z * sZ;
    B += blockIdx.y * sB + blockIdx.z * sZ;
    C += blockIdx.y * sC + blockIdx.z * sZ;
    __shared__ float As[2][KC][BM + 4];
    __shared__ float Bs[2][KC][BM + 4];
    const int64_t bid = (tiles == 2 || tiles == 4) ? (int64_t)blockIdx.x : xcd_remap32(blockIdx.x, gridDim.x);
    int64_t ti, tj;
    if (tiles == 0 || tiles == 3 || tiles == 4) {
        ti = bid % tilesM;
        tj = bid / tilesM;
    } else {
        lower_tile32(bid, ti, tj);
    }
    const int64_t row0 = ti * BM, col0 = tj * BM;
    const int64_t kbeg = (tiles == 2) ? ti * BM : (tiles == 4) ? tj * BM : 0;
    const int64_t kend = (tiles == 3 && (ti + 1) * BM < K) ? (ti + 1) * BM : K;
    const int64_t nch = kend > kbeg ? (kend - kbeg) / KC : 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l32 = lane & 31, h = lane >> 5;
    const int wr = wave / WCn, wc = wave % WCn;

    f16v acc[NBR][NBC];
    double acc64[FR][FC][FOLD > 0 ? 16 : 1];
#pragma unroll
    for (int i = 0; i < NBR; ++i)
#pragma unroll
        for (int j = 0; j < NBC; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
    if constexpr (FOLD > 0) {
#pragma unroll
        for (int i = 0; i < FR; ++i)
#pragma unroll
            for (int j = 0; j < FC; ++j)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc64[i][j][v] = 0.0;
    }

    float4 ra[AL], rb[AL];
    if (nch > 0) {
        g32_load<BM, !TA, NT, KC>(A, lda, row0, M, kbeg, ra);
        g32_load<BM, TB, NT, KC>(B, ldb, col0, N, kbeg, rb);
        g32_store<BM, !TA, NT, KC>(As[0], ra);
        g32_store<BM, TB, NT, KC>(Bs[0], rb);
    }
    __syncthreads();
    for (int64_t c = 0; c < nch; ++c) {
        const int buf = (int)(c & 1);
        const bool more = c + 1 < nch;
        if (more) {
            g32_load<BM, !TA, NT, KC>(A, lda, row0, M, kbeg + (c + 1) * KC, ra);
            g32_load<BM, TB, NT, KC>(B, ldb, col0, N, kbeg + (c + 1) * KC, rb);
        }
#pragma unroll
        for (int s = 0; s < KC / 2; ++s) {
            const int kk = (KC / 2) * h + s;
            float a[NBR], b[NBC];
#pragma unroll
            for (int i = 0; i < NBR; ++i) a[i] = As[buf][kk][wr * WTR + i * 32 + l32];
#pragma unroll
            for (int j = 0; j < NBC; ++j) b[j] = Bs[buf][kk][wc * WTC + j * 32 + l32];
#pragma unroll
            for (int i = 0; i < NBR; ++i)
#pragma unroll
                for (int j = 0; j < NBC; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if constexpr (FOLD > 0) {
            if ((c % FOLD) == FOLD - 1 || !more) {  // block-uniform branch
#pragma unroll
                for (int i = 0; i < NBR; ++i)
#pragma unroll
                    for (int j = 0; j < NBC; ++j)
#pragma unroll
                        for (int v = 0; v < 16; ++v) {
                            acc64[i][j][v] += (double)acc[i][j][v];
                            acc[i][j][v] = 0.f;
                        }
            }
        }
        if (more) {
            g32_store<BM, !TA, NT, KC>(As[buf ^ 1], ra);
            g32_store<BM, TB, NT, KC>(Bs[buf ^ 1], rb);
        }
        __syncthreads();
    }
    // v_mfma_f32_32x32x2f32 D layout: item v -> row 8*(v/4) + 4*(lane/32) + v%4, column lane%32,
    // so items 4g..4g+3 of a lane are four consecutive rows: one float4 per group
#pragma unroll
    for (int i = 0; i < NBR; ++i)
#pragma unroll
        for (int j = 0; j < NBC; ++j) {
            const int64_t rb0 = row0 + wr * WTR + i * 32, cb0 = col0 + wc * WTC + j * 32;
            if (rb0 >= M || cb0 >= N) continue;
            float* p = C + rb0 + 4 * h + (cb0 + l32) * ldc;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                float4* q = reinterpret_cast<float4*>(p + 8 * g);
                float4 cv = (beta == 0.f) ? float4{0.f, 0.f, 0.f, 0.f} : *q;
                float o[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int v = 4 * g + e;
                    const float cvi = (e == 0) ? cv.x : (e == 1) ? cv.y : (e == 2) ? cv.z : cv.w;
                    if constexpr (FOLD > 0)
                        o[e] = (float)((double)beta * (double)cvi + (double)alpha * acc64[i][j][v]);
                    else
                        o[e] = beta * cvi + alpha * acc[i][j][v];
                }
                *q = float4{o[0], o[1], o[2], o[3]};
            }
        }
}

template <int BM, int FOLD, int NW, int KC = 16>
static void gemm32_launch(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, float alpha,
                          const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C,
                          int64_t ldc, int tiles, int batch, int64_t sA, int64_t sB, int64_t sC, int nz,
                          int64_t sZ) {
    const int64_t tm = (M + BM - 1) / BM, tn = (N + BM - 1) / BM;
    const int64_t nblk = (tiles == 0 || tiles == 3 || tiles == 4) ? tm * tn : tm * (tm + 1) / 2;
    dim3 grid((unsigned)nblk, (unsigned)batch, (unsigned)nz), block(64 * NW);
#define GPL_G32(TAv, TBv)                                                                                        \
    hipLaunchKernelGGL((k_gemm32<TAv, TBv, BM, FOLD, NW, KC>), grid, block, 0, s, M, N, K, tm, alpha, A, lda, B, ldb, \
                       beta, C, ldc, tiles, sA, sB, sC, sZ)
    if (!ta && !tb) GPL_G32(0, 0);
    else if (!ta && tb) GPL_G32(0, 1);
    else if (ta && !tb) GPL_G32(1, 0);
    else GPL_G32(1, 1);
#undef GPL_G32
}

// fold != 0: FP32 sums folded into FP64 every 64 k; 128-wide tiles use eight
// waves there (64x32 per wave keeps the FP64 accumulators at four waves/SIMD).
// fold == 0: plain FP32 accumulation.  bm == 0 picks 128-wide tiles when the
// grid holds at least 256 of them.
void gemm_f32_batched2(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, float alpha, const float* A,
                       int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int tiles, int bm,
                       int batch, int64_t sA, int64_t sB, int64_t sC, int nz, int64_t sZ, int fold) {
    if (M <= 0 || N <= 0 || batch <= 0 || nz <= 0) return;
    if (bm == 0) {
        const int64_t tm = (M + 127) / 128, tn = (N + 127) / 128;
        const int64_t n128 = ((tiles == 0 || tiles == 3 || tiles == 4) ? tm * tn : tm * (tm + 1) / 2) * batch * nz;
        bm = (n128 >= 256) ? 128 : 64;
    }
#define GPL_G32L(BMv, FOLDv, NWv)                                                                                     \
    gemm32_launch<BMv, FOLDv, NWv>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles, batch, sA, sB, sC, \
                                   nz, sZ)
    if (fold) {
        if (bm == 128) GPL_G32L(128, 4, 8);
        else GPL_G32L(64, 4, 4);
    } else {
        if (bm == 128) GPL_G32L(128, 0, 4);
        else GPL_G32L(64, 0, 4);
    }
#undef GPL_G32L
}

void gemm_f32_batched(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, float alpha, const float* A,
                      int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int tiles, int bm,
                      int batch, int64_t sA, int64_t sB, int64_t sC, int fold) {
    gemm_f32_batched2(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles, bm, batch, sA, sB, sC, 1, 0,
                      fold);
}

void gemm_f32(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, float alpha, const float* A,
              int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int tiles, int bm, int fold) {
    gemm_f32_batched(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles, bm, 1, 0, 0, 0, fold);
}

// ---------------------------------------------------------------------------
// FP32-matrix helpers with FP64 vectors
// ---------------------------------------------------------------------------
// A = I over all T x T tiles (the 128-wide triangular GEMM modes read the tile
// just above each diagonal tile, so upper tiles must hold zeros)
__global__ void k_set_identity32(float* __restrict__ A, int64_t ld, int64_t T, int64_t sA) {
    A += blockIdx.y * sA;
    const int64_t ti = blockIdx.x % T, tj = blockIdx.x / T;
    const int t = threadIdx.x, i = t & 63;
    for (int c = t >> 6; c < 64; c += 4) {
        const int64_t gi = ti * 64 + i, gj = tj * 64 + c;
        A[gi + gj * ld] = (gi == gj) ? 1.f : 0.f;
    }
}

void set_identity_f32_batched(hipStream_t s, float* A, int64_t ld, int64_t N, int batch, int64_t sA) {
    const int64_t T = N / 64;
    hipLaunchKernelGGL(k_set_identity32, dim3((unsigned)(T * T), (unsigned)batch), dim3(256), 0, s, A, ld, T, sA);
}

void set_identity_f32(hipStream_t s, float* A, int64_t ld, int64_t N) {
    set_identity_f32_batched(s, A, ld, N, 1, 0);
}

// Column j = c0 + blockIdx.x of X (FP32, N x N, ld N, batch stride sX): zero
// above its 64x64 diagonal block, the block's column of Dinv (FP64, 64x64
// column-major per block, zero above its diagonal, batch stride sD) rounded to
// FP32 on it; the rows below are left to the caller.
__global__ __launch_bounds__(256) void k_trtri_leaves32(float* __restrict__ X, int64_t N,
                                                        const double* __restrict__ Dinv, int64_t sX, int64_t sD,
                                                        int64_t c0) {
    X += blockIdx.y * sX;
    Dinv += blockIdx.y * sD;
    const int64_t j = c0 + blockIdx.x, kb = j / 64, r0 = kb * 64;
    const double* D = Dinv + kb * 4096 + (j - r0) * 64;
    for (int64_t i = threadIdx.x; i < r0 + 64; i += 256) X[i + j * N] = (i < r0) ? 0.f : (float)D[i - r0];
}

void trtri_leaves_f32_batched(hipStream_t s, float* X, int64_t N, const double* Dinv, int batch, int64_t sX,
                              int64_t sD, int64_t c0, int64_t ncols) {
    if (ncols <= 0) return;
    hipLaunchKernelGGL(k_trtri_leaves32, dim3((unsigned)ncols, (unsigned)batch), dim3(256), 0, s, X, N, Dinv, sX, sD,
                       c0);
}

// y = X b (trans 0) or X^T b (trans 1), X lower triangular FP32 (64x64 tiles,
// column-major), vectors FP64.  Grid (T, nchunk, batch): block (ti, ch) sums
// the tiles of output block ti that fall in chunk ch; for X^T the stored tile
// (r, ti) is staged through LDS so every global read is coalesced.  The chunk
// partials (part: batch * nchunk * N) are added in a fixed order afterwards.
template <typename TX>
__global__ __launch_bounds__(256) void k_trmv32(const TX* __restrict__ X, int64_t ld, int64_t T, int nchunk,
                                                const double* __restrict__ b, double* __restrict__ part, int trans,
                                                int64_t sX, int64_t sv) {
    __shared__ TX tile[64][65];
    __shared__ double bc[64];
    __shared__ double red[256];
    X += blockIdx.z * sX;
    b += blockIdx.z * sv;
    part += blockIdx.z * (int64_t)nchunk * T * 64;
    const int64_t ti = blockIdx.x;
    const int ch = blockIdx.y;
    const int t = threadIdx.x, e = t & 63, p = t >> 6;
    // tiles of this output block: tj in [0, ti] (X b) or r in [ti, T) (X^T b), split into nchunk ranges
    const int64_t lo = trans ? ti : 0, hi = trans ? T : ti + 1, cnt = hi - lo;
    const int64_t per = (cnt + nchunk - 1) / nchunk, tb = lo + ch * per, te = (tb + per < hi) ? tb + per : hi;
    double s = 0.0;
    for (int64_t q = tb; q < te; ++q) {
        __syncthreads();
        if (t < 64) bc[t] = b[q * 64 + t];
        if (trans) {
            const TX* Xt = X + q * 64 + ti * 64 * ld;  // stored tile (q, ti)
#pragma unroll 4
            for (int u = 0; u < 16; ++u) {
                const int c = p * 16 + u;
                tile[c][e] = Xt[e + (int64_t)c * ld];
            }
        }
        __syncthreads();
        if (!trans) {
            const TX* Xt = X + ti * 64 + q * 64 * ld;
#pragma unroll 4
            for (int u = 0; u < 16; ++u) {
                const int c = p * 16 + u;
                s += (double)Xt[e + (int64_t)c * ld] * bc[c];
            }
        } else {
#pragma unroll 4
            for (int u = 0; u < 16; ++u) {
                const int r = p * 16 + u;
                s += (double)tile[e][r] * bc[r];
            }
        }
    }
    red[t] = s;
    __syncthreads();
    if (t < 64) part[(int64_t)ch * T * 64 + ti * 64 + t] = red[t] + red[t + 64] + red[t + 128] + red[t + 192];
}

void trmv_lower_f32_batched(hipStream_t s, const float* X, int64_t ld, int64_t N, const double* b, double* y,
                            int trans, int batch, int64_t sX, int64_t sv, double* part, int nchunk) {
    const int64_t T = N / 64;
    hipLaunchKernelGGL(k_trmv32<float>, dim3((unsigned)T, (unsigned)nchunk, (unsigned)batch), dim3(256), 0, s, X, ld,
                       T, nchunk, b, part, trans, sX, sv);
    chunk_sum_f64(s, part, N, nchunk, y, batch, sv);
}

// the same over an FP64 matrix (one problem)
void trmv_lower_f64_chunked(hipStream_t s, const double* X, int64_t ld, int64_t N, const double* b, double* y,
                            int trans, double* part, int nchunk) {
    const int64_t T = N / 64;
    hipLaunchKernelGGL(k_trmv32<double>, dim3((unsigned)T, (unsigned)nchunk, 1u), dim3(256), 0, s, X, ld, T, nchunk,
                       b, part, trans, (int64_t)0, (int64_t)0);
    chunk_sum_f64(s, part, N, nchunk, y, 1, 0);
}

void trmv_lower_f32(hipStream_t s, const float* X, int64_t ld, int64_t N, const double* b, double* y, int trans,
                    double* part, int nchunk) {
    trmv_lower_f32_batched(s, X, ld, N, b, y, trans, 1, 0, 0, part, nchunk);
}

__global__ __launch_bounds__(256) void k_colsumsq32(const float* __restrict__ A, int64_t ld, int64_t nrow,
                                                    int64_t ncol, double* __restrict__ out, int lower, int64_t sA,
                                                    int64_t sO) {
    A += blockIdx.y * sA;
    out += blockIdx.y * sO;
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= ncol) return;
    const float* col = A + j * ld;
    double s = 0.0;
    for (int64_t i = (lower ? (j / 64) * 64 : 0) + lane; i < nrow; i += 64) {
        const double v = col[i];
        s += v * v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) out[j] = s;
}

void colsumsq_f32_batched(hipStream_t s, const float* A, int64_t ld, int64_t nrow, int64_t ncol, double* out,
                          int lower, int batch, int64_t sA, int64_t sO) {
    if (ncol <= 0) return;
    hipLaunchKernelGGL(k_colsumsq32, dim3((unsigned)((ncol + 3) / 4), (unsigned)batch), dim3(256), 0, s, A, ld, nrow,
                       ncol, out, lower, sA, sO);
}

void colsumsq_f32(hipStream_t s, const float* A, int64_t ld, int64_t nrow, int64_t ncol, double* out, int lower) {
    colsumsq_f32_batched(s, A, ld, nrow, ncol, out, lower, 1, 0, 0);
}

// max |d| / max |u| (one workgroup): the refinement stopping test
__global__ __launch_bounds__(1024) void k_relmax(const double* __restrict__ d, const double* __restrict__ u,
                                                 int64_t n, double* __restrict__ out) {
    __shared__ double rd[1024], ru[1024];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int64_t i = t; i < n; i += 1024) {
        a = fmax(a, fabs(d[i]));
        b = fmax(b, fabs(u[i]));
    }
    rd[t] = a;
    ru[t] = b;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o) {
            rd[t] = fmax(rd[t], rd[t + o]);
            ru[t] = fmax(ru[t], ru[t + o]);
        }
        __syncthreads();
    }
    if (t == 0) out[0] = (ru[0] > 0.0) ? rd[0] / ru[0] : rd[0];
}

void rel_maxabs(hipStream_t s, const double* d, const double* u, int64_t n, double* out) {
    hipLaunchKernelGGL(k_relmax, dim3(1), dim3(1024), 0, s, d, u, n, out);
}

}  // namespace gpl
