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
// LDS-staged GEMM: BM x BM output tile per 256-thread workgroup (BM = 128 or 64),
// 16-wide k chunks double-buffered in LDS as [k][m] (global -> registers for
// chunk c+1 is issued before the MFMAs of chunk c and written to LDS after
// them).  Each wave owns a (BM/2)^2 quadrant of v_mfma_f64_16x16x4 blocks.
// M and N are multiples of 64; a 128 tile whose second half falls outside is
// masked (zero loads, skipped stores).
// ---------------------------------------------------------------------------
template <typename T> struct vec2_of;
template <> struct vec2_of<double> { typedef double2 type; };
template <> struct vec2_of<float> { typedef float2 type; };

// LDS tiles are [k][BM + 1]: the odd row stride keeps the eight lanes that
// write one column's k values on distinct banks.
template <int BM, int MCONT, typename T, int NT = 256>
__device__ __forceinline__ void lds_gemm_load(const T* P, int64_t ld, int64_t base, int64_t lim, int64_t k0,
                                              double2 (&r)[BM * 8 / NT]) {
    typedef typename vec2_of<T>::type V2;
    constexpr int AL = BM * 8 / NT;
    const int t = threadIdx.x;
    if (MCONT) {  // element (m, k) at P[m + k*ld]: pairs along m, 128-byte rows per 8 lanes
        const int mp = (t % (BM / 2)) * 2, kq = t / (BM / 2), ks = NT / (BM / 2);
        const bool ok = base + mp < lim;
#pragma unroll
        for (int j = 0; j < AL; ++j) {
            const int64_t k = k0 + kq + ks * j;
            if (ok) {
                const V2 v = *reinterpret_cast<const V2*>(P + base + mp + k * ld);
                r[j] = double2{(double)v.x, (double)v.y};
            } else {
                r[j] = double2{0.0, 0.0};
            }
        }
    } else {  // element (m, k) at P[k + m*ld]: 8 lanes read one column's 16 k contiguously
        const int kq = t & 7, m = t >> 3;
#pragma unroll
        for (int j = 0; j < AL; ++j) {
            const int64_t mm = base + m + (NT / 8) * j;
            if (mm < lim) {
                const V2 v = *reinterpret_cast<const V2*>(P + (k0 + 2 * kq) + mm * ld);
                r[j] = double2{(double)v.x, (double)v.y};
            } else {
                r[j] = double2{0.0, 0.0};
            }
        }
    }
}

template <int BM, int MCONT, int NT = 256>
__device__ __forceinline__ void lds_gemm_store(double (*S)[BM + 1], const double2 (&r)[BM * 8 / NT]) {
    constexpr int AL = BM * 8 / NT;
    const int t = threadIdx.x;
    if (MCONT) {
        const int mp = (t % (BM / 2)) * 2, kq = t / (BM / 2), ks = NT / (BM / 2);
#pragma unroll
        for (int j = 0; j < AL; ++j) {
            S[kq + ks * j][mp] = r[j].x;
            S[kq + ks * j][mp + 1] = r[j].y;
        }
    } else {
        const int kq = t & 7, m = t >> 3;
#pragma unroll
        for (int j = 0; j < AL; ++j) {
            S[2 * kq][m + (NT / 8) * j] = r[j].x;
            S[2 * kq + 1][m + (NT / 8) * j] = r[j].y;
        }
    }
}

// C = beta*C + alpha*op(A) op(B), FP64 MFMA; A, B and C may be stored as float or double.
// NW waves in a 2 x NW/2 grid each own a (BM/2) x (2BM/NW) block of 16x16 MFMA tiles.
template <int TA, int TB, int BM, typename TAs, typename TBs, typename TCs, int NW = 4>
__global__ __launch_bounds__(64 * NW) void k_gemm_lds(int64_t M, int64_t N, int64_t K, int64_t tilesM, double alpha,
                                                      const TAs* A, int64_t lda, const TBs* B, int64_t ldb,
                                                      double beta, TCs* C, int64_t ldc, int tiles, int64_t sA,
                                                      int64_t sB, int64_t sC) {
    constexpr int NT = 64 * NW, WCn = NW / 2, WTR = BM / 2, WTC = BM / WCn;
    constexpr int NBR = WTR / 16, NBC = WTC / 16, AL = BM * 8 / NT;
    static_assert(NBR >= 1 && NBC >= 1 && AL >= 1, "tile shape");
    A += blockIdx.y * sA;  // batch: one problem per grid row
    B += blockIdx.y * sB;
    C += blockIdx.y * sC;
    __shared__ double As[2][16][BM + 1];
    __shared__ double Bs[2][16][BM + 1];
    // Contiguous tile ranges per XCD share operand panels in L2; with a
    // k range that depends on the tile row (tiles == 2) or column (tiles == 4,
    // whose columns are contiguous in the tile order) work per tile varies
    // widely, so the hardware's round-robin placement is kept there to balance
    // the XCDs.
    const int64_t bid = (tiles == 2 || tiles == 4) ? (int64_t)blockIdx.x : xcd_remap(blockIdx.x, gridDim.x);
    int64_t ti, tj;
    if (tiles == 0 || tiles == 3 || tiles == 4) {
        ti = bid % tilesM;
        tj = bid / tilesM;
    } else {
        lower_tile(bid, ti, tj);
    }
    const int64_t row0 = ti * BM, col0 = tj * BM;
    const int64_t kbeg = (tiles == 2) ? ti * BM : (tiles == 4) ? tj * BM : 0;
    const int64_t kend = (tiles == 3 && (ti + 1) * BM < K) ? (ti + 1) * BM : K;
    const int64_t nch = kend > kbeg ? (kend - kbeg) / 16 : 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int wr = wave / WCn, wc = wave % WCn;

    d4 acc[NBR][NBC];
#pragma unroll
    for (int i = 0; i < NBR; ++i)
#pragma unroll
        for (int j = 0; j < NBC; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    auto chunk = [&](int buf) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int kk = 4 * g + s;
            double a[NBR], b[NBC];
#pragma unroll
            for (int i = 0; i < NBR; ++i) a[i] = As[buf][kk][wr * WTR + i * 16 + li];
#pragma unroll
            for (int j = 0; j < NBC; ++j) b[j] = Bs[buf][kk][wc * WTC + j * 16 + li];
#pragma unroll
            for (int i = 0; i < NBR; ++i)
#pragma unroll
                for (int j = 0; j < NBC; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    };
    double2 ra[AL], rb[AL];
    if (nch > 0) {
        lds_gemm_load<BM, !TA, TAs, NT>(A, lda, row0, M, kbeg, ra);
        lds_gemm_load<BM, TB, TBs, NT>(B, ldb, col0, N, kbeg, rb);
        lds_gemm_store<BM, !TA, NT>(As[0], ra);
        lds_gemm_store<BM, TB, NT>(Bs[0], rb);
    }
    __syncthreads();
    for (int64_t c = 0; c < nch; ++c) {
        const int buf = (int)(c & 1);
        const bool more = c + 1 < nch;
        if (more) {
            lds_gemm_load<BM, !TA, TAs, NT>(A, lda, row0, M, kbeg + (c + 1) * 16, ra);
            lds_gemm_load<BM, TB, TBs, NT>(B, ldb, col0, N, kbeg + (c + 1) * 16, rb);
        }
        chunk(buf);
        if (more) {
            lds_gemm_store<BM, !TA, NT>(As[buf ^ 1], ra);
            lds_gemm_store<BM, TB, NT>(Bs[buf ^ 1], rb);
        }
        __syncthreads();
    }
    // v_mfma_f64_16x16x4 D layout: lane holds rows g, g+4, g+8, g+12 of column li.
    // With beta != 0 every C value is loaded before the first store (a load
    // placed after a store that may alias it waits for that store's round
    // trip).  The eight-wave 128 tile pays for the registers with one wave per
    // SIMD less and still comes out ahead (Cholesky trailing updates at
    // n = 4096 and 8192); the FP32 GEMM, whose 128 tile measured slower with
    // the same change, keeps the interleaved order.
    TCs cv[NBR][NBC][4];
    if (beta != 0.0) {
#pragma unroll
        for (int i = 0; i < NBR; ++i)
#pragma unroll
            for (int j = 0; j < NBC; ++j) {
                const int64_t rb0 = row0 + wr * WTR + i * 16, cb0 = col0 + wc * WTC + j * 16;
                if (rb0 >= M || cb0 >= N) continue;
                const TCs* p = C + (rb0 + g) + (cb0 + li) * ldc;
#pragma unroll
                for (int v = 0; v < 4; ++v) cv[i][j][v] = p[4 * v];
            }
    }
#pragma unroll
    for (int i = 0; i < NBR; ++i)
#pragma unroll
        for (int j = 0; j < NBC; ++j) {
            const int64_t rb0 = row0 + wr * WTR + i * 16, cb0 = col0 + wc * WTC + j * 16;
            if (rb0 >= M || cb0 >= N) continue;
            TCs* p = C + (rb0 + g) + (cb0 + li) * ldc;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double o = alpha * acc[i][j][v];
                p[4 * v] = (TCs)((beta == 0.0) ? o : beta * (double)cv[i][j][v] + o);
            }
        }
}

template <int BM, typename TAs, typename TBs, typename TCs, int NW = 4>
static void gemm_lds_launch(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                            const TAs* A, int64_t lda, const TBs* B, int64_t ldb, double beta, TCs* C,
                            int64_t ldc, int tiles, int batch = 1, int64_t sA = 0, int64_t sB = 0,
                            int64_t sC = 0) {
    const int64_t tm = (M + BM - 1) / BM, tn = (N + BM - 1) / BM;
    const int64_t nblk = (tiles == 0 || tiles == 3 || tiles == 4) ? tm * tn : tm * (tm + 1) / 2;
    dim3 grid((unsigned)nblk, (unsigned)batch), block(64 * NW);
#define GPL_GL(TAv, TBv)                                                                                      \
    hipLaunchKernelGGL((k_gemm_lds<TAv, TBv, BM, TAs, TBs, TCs, NW>), grid, block, 0, s, M, N, K, tm, alpha, A, \
                       lda, B, ldb, beta, C, ldc, tiles, sA, sB, sC)
    if (!ta && !tb) GPL_GL(0, 0);
    else if (!ta && tb) GPL_GL(0, 1);
    else if (ta && !tb) GPL_GL(1, 0);
    else GPL_GL(1, 1);
#undef GPL_GL
}

// Tile choice for FP64 storage.  Short inner dimensions (the panel and trailing
// updates, k of 64 to 256) are bound by the chunk loop's load latency and want
// the most workgroups: 64x64 tiles on four waves.  Long inner dimensions (B^-1 =
// X^T X, prediction's L^-1 K_t^T, k = n) run faster on 128x128 tiles with eight
// waves (half the operand reads per flop at the same occupancy): 57 against 46
// TFLOP/s for X^T X at n = 8192.  bm = 64 or 128 forces the 64/4 or 128/8 form.
static int pick_bm(int64_t M, int64_t N, int64_t K, int tiles, int batch = 1) {
    const int64_t tm = (M + 127) / 128, tn = (N + 127) / 128;
    const int64_t n128 = (tiles == 0 || tiles == 3 || tiles == 4) ? tm * tn : tm * (tm + 1) / 2;
    return (K >= 1024 && n128 * batch >= 256) ? 128 : 64;
}

void gemm2_f64(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
               int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int tiles, int bm) {
    if (M <= 0 || N <= 0) return;
    if (bm == 0) bm = pick_bm(M, N, K, tiles);
    if (bm == 128)
        gemm_lds_launch<128, double, double, double, 8>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles);
    else gemm_lds_launch<64>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles);
}

// `batch` independent FP64 GEMMs (strides in elements), one grid row per problem
void gemm2_f64_batched(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
                       int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int tiles,
                       int batch, int64_t sA, int64_t sB, int64_t sC) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    if (pick_bm(M, N, K, tiles, batch) == 128)
        gemm_lds_launch<128, double, double, double, 8>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                                        tiles, batch, sA, sB, sC);
    else gemm_lds_launch<64>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles, batch, sA, sB, sC);
}

// out[i] = sum over p < parts of W[p * stride + i], in order of p (deterministic)
__global__ void k_sum_parts(const double* __restrict__ W, int64_t stride, int parts, int64_t n,
                            double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = W[i];
    for (int p = 1; p < parts; ++p) acc += W[p * stride + i];
    out[i] = acc;
}

void sum_parts_f64(hipStream_t s, const double* W, int64_t stride, int parts, int64_t n, double* out) {
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, W, stride, parts, n, out);
}

// FP64-compute GEMM on FP32 storage for A/C with an FP64 B (panel solves of the mixed path)
void gemm_mixed_f32f64(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha, const float* A,
                       int64_t lda, const double* B, int64_t ldb, double beta, float* C, int64_t ldc, int tiles) {
    if (M <= 0 || N <= 0) return;
    gemm_lds_launch<64>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles);
}

// FP64 B on the left: C(float) = alpha * A(double) op B(float)
void gemm_mixed_f64f32(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                       const double* A, int64_t lda, const float* B, int64_t ldb, double beta, float* C, int64_t ldc,
                       int tiles) {
    if (M <= 0 || N <= 0) return;
    gemm_lds_launch<64>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles);
}

void gemm_mixed_f32f64_batched(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                               const float* A, int64_t lda, const double* B, int64_t ldb, double beta, float* C,
                               int64_t ldc, int tiles, int batch, int64_t sA, int64_t sB, int64_t sC) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    gemm_lds_launch<64>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles, batch, sA, sB, sC);
}

void gemm_mixed_f64f32_batched(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                               const double* A, int64_t lda, const float* B, int64_t ldb, double beta, float* C,
                               int64_t ldc, int tiles, int batch, int64_t sA, int64_t sB, int64_t sC) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    gemm_lds_launch<64>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles, batch, sA, sB, sC);
}

void gemm_mixed_f32f64f64(hipStream_t s, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                          const float* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                          int64_t ldc, int tiles) {
    if (M <= 0 || N <= 0) return;
    if (pick_bm(M, N, K, tiles) == 128)
        gemm_lds_launch<128, float, double, double, 8>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles);
    else gemm_lds_launch<64>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles);
}

// ---------------------------------------------------------------------------
// Diagonal block (64x64), FP64 arithmetic on FP64 or FP32 storage.
// Thread (row i = t%64, wave w = t/64) keeps elements (i, 16w+q), q < 16, in
// registers for the whole factorisation, which runs in four phases of 16
// columns (one wave each, exchanging pivots by readlane) with a rank-16 update
// of the waves to the right between phases.  The inverse
// (written to Dinv, FP64) is then formed by 16x16 blocks: DPP forward
// substitution inside the diagonal blocks, f64 MFMA for the block rows.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 1/sqrt(d) for d > 0: hardware estimate refined by two Newton steps
// y <- y (3 - d y^2) / 2 (each roughly doubles the correct bits), a short
// dependent chain compared with a correctly rounded sqrt followed by a divide.
// Non-positive or NaN d propagates NaN, which the pivot check reports.
__device__ __forceinline__ double rsqrt_f64(double d) {
    double y = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

template <int KQ>
__device__ __forceinline__ double quad_bcast_f64(double v) {
    constexpr int ctrl = KQ | (KQ << 2) | (KQ << 4) | (KQ << 6);  // quad_perm [KQ,KQ,KQ,KQ]
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), ctrl, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One 64x64 diagonal block (block k of the N x N matrix A, leading dimension
// ld) on one 256-thread workgroup: factor in place, Dinv + 4096k = the block's
// inverse (when want_inv), ldiag[64k + i] = log L_ii, info = first failing row
// + 1.  Dsrc (optional): the block's FP64 values (64x64 column-major) to factor
// in place of A's.  IX: the block's values (FP64) are already in Xs
// (Xs[col][row]), in place of either.  Lp, Xs (64 x 65), rinv, piv (64): workgroup LDS;
// Lp[col][row] = L(row, col) (strictly lower after the factor), Xs[col][row] =
// (L_kk^-1)(row, col).  MU (pcol: one LDS int): the waves to the right of
// phase p apply its columns as rank-1 updates while wave p factors, following
// a column counter in LDS, instead of one rank-16 update after a barrier.
template <typename T, bool MU, bool IX = false>
__device__ __forceinline__ void diag_block64(T* __restrict__ A, int64_t ld, int64_t k, double* __restrict__ Dinv,
                                             double* __restrict__ ldiag, int* __restrict__ info, int want_inv,
                                             const double* __restrict__ Dsrc, double (*Lp)[65], double (*Xs)[65],
                                             double* rinv, double* piv, int* pcol = nullptr) {
    // One 64x64 diagonal block per workgroup, FP64 arithmetic.  Wave w holds
    // columns 16w..16w+15 of every row (lane = row).  Factor: four phases; in
    // phase p wave p factors its 16 columns alone (pivots and multipliers are
    // exchanged with readlane, no barrier), publishes them to LDS, and the
    // waves to its right apply the rank-16 update.  Inverse: by 16x16 blocks
    // (below).
    T* blk = A + k * 64 * (ld + 1);
    const int t = threadIdx.x, i = t & 63, w = t >> 6;
    if (MU) {
        if (t == 0) *pcol = 0;
        __syncthreads();
    }
    double a[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = 16 * w + q;
        a[q] = (c <= i) ? (IX ? Xs[c][i] : Dsrc ? Dsrc[i + c * 64] : (double)blk[i + (int64_t)c * ld]) : 0.0;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        if (w == p) {
            double myinv = 0.0, mypiv = 0.0;
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const int j = 16 * p + jj;
                const double d = readlane_f64(a[jj], j);
                const double inv = rsqrt_f64(d);
                const double rt = d * inv;
                const double l = (i > j) ? a[jj] * inv : ((i == j) ? rt : 0.0);
                a[jj] = l;
                myinv = (i == j) ? inv : myinv;
                mypiv = (i == j) ? d : mypiv;
                // rows above a column are cleared when that column is reached.
                // Column j goes to LDS, which is also where the multipliers
                // L(16p + cc, j) of the later columns are read from (one
                // address for the whole wave); only the next pivot's column
                // takes its multiplier by readlane, off the LDS round trip.
                Lp[j][i] = l;
                if (jj + 1 < 16) a[jj + 1] -= l * readlane_f64(l, j + 1);
#pragma unroll
                for (int cc = jj + 2; cc < 16; ++cc) a[cc] -= l * Lp[j][16 * p + cc];
                // MU: every four columns, tell the waves to the right how many
                // columns of L are in LDS (release: the columns land first)
                if (MU && (jj & 3) == 3)
                    __hip_atomic_store(pcol, j + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            if (i >= 16 * p && i < 16 * p + 16) {
                rinv[i] = myinv;
                piv[i] = mypiv;
            }
        } else if (MU && w > p) {
            // Pipelined: apply the rank-1 updates of phase p's columns to this
            // wave's columns as they are published, four at a time, while wave p
            // factors the next ones; no barrier between the phases, and the wave
            // of the next phase starts right after the last four.  (Groups
            // shrinking to single columns at the end of the phase, applied by a
            // rolled loop, measured slower: 31 against 24 us per 64 columns.)
#pragma unroll 1
            for (int g = 0; g < 4; ++g) {
                const int need = 16 * p + 4 * g + 4;
                while (__hip_atomic_load(pcol, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < need)
                    __builtin_amdgcn_s_sleep(1);
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const int j = need - 4 + jj;
                    const double li = Lp[j][i];
#pragma unroll
                    for (int q = 0; q < 16; ++q) a[q] -= li * Lp[j][16 * w + q];
                }
            }
        }
        if (!MU) __syncthreads();
        if (!MU && w > p) {
            double li[16];
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) li[kk] = Lp[16 * p + kk][i];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int col = 16 * w + q;
                double sum = 0.0;
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) sum += li[kk] * Lp[16 * p + kk][col];
                if (col <= i) a[q] -= sum;
            }
        }
    }
    if (MU) __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = 16 * w + q;
        blk[i + (int64_t)c * ld] = (T)((c <= i) ? a[q] : 0.0);
    }
    if (t < 64) Lp[t][t] = 0.0;  // the substitution uses the strictly lower part
    __syncthreads();
    if (t < 64) {
        ldiag[k * 64 + t] = 0.5 * log(piv[t]);
        const unsigned long long bad = __ballot(!(piv[t] > 0.0));
        if (t == 0 && bad) atomicCAS(info, 0, (int)(k * 64 + __ffsll((long long)bad)));
    }
    if (!want_inv) return;
    // Inverse by 16x16 blocks.  (1) Each wave inverts its diagonal block by
    // forward substitution, four lanes per column exchanging values by DPP (16
    // dependent steps).  (2) Block rows bi = 1..3 in turn, wave j < bi forming
    // X_{bi,j} = -X_{bi,bi} T, T = sum_{j <= kb < bi} L_{bi,kb} X_{kb,j}, on
    // v_mfma_f64_16x16x4.  T is left in the D layout (lane (li, g) holds rows
    // g + 4v of column li), which is the A-operand layout of T^T with k = g + 4v,
    // so the product is taken as Y^T = T^T X_{bi,bi}^T without leaving registers.
    const int lane = t & 63, li = lane & 15, g = lane >> 4;
    {
        const int c = lane >> 2, q = lane & 3, base = 16 * w;
        double b[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) b[r] = (4 * r + q == c) ? 1.0 : 0.0;
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            const int kr = kk >> 2;
            double xk = b[kr] * rinv[base + kk];
            switch (kk & 3) {
                case 0: xk = quad_bcast_f64<0>(xk); break;
                case 1: xk = quad_bcast_f64<1>(xk); break;
                case 2: xk = quad_bcast_f64<2>(xk); break;
                default: xk = quad_bcast_f64<3>(xk); break;
            }
            if (q == (kk & 3)) b[kr] = xk;
#pragma unroll
            for (int r = 0; r < 4; ++r) b[r] -= Lp[base + kk][base + 4 * r + q] * xk;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) Xs[base + c][base + 4 * r + q] = b[r];
        for (int r = q; r < base; r += 4) Xs[base + c][r] = 0.0;  // above the diagonal block
    }
    __syncthreads();
#pragma unroll
    for (int bi = 1; bi < 4; ++bi) {
        if (w < bi) {
            const int r0 = 16 * bi, c0 = 16 * w;
            d4 acc = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kb = 0; kb < r0; kb += 4)  // X is zero above its diagonal blocks
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Lp[kb + g][r0 + li], Xs[c0 + li][kb + g], acc, 0, 0, 0);
            d4 y = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int v = 0; v < 4; ++v)
                y = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[v], Xs[r0 + g + 4 * v][r0 + li], y, 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 4; ++v) Xs[c0 + g + 4 * v][r0 + li] = -y[v];
        }
        __syncthreads();
    }
    double* D = Dinv + k * 64 * 64;
    for (int e = t; e < 64 * 64; e += 256) D[(e & 63) + (e >> 6) * 64] = Xs[e >> 6][e & 63];
}

template <typename T>
__global__ __launch_bounds__(256) void k_potrf_diag3(T* __restrict__ A, int64_t ld, int64_t k,
                                                     double* __restrict__ Dinv, double* __restrict__ ldiag,
                                                     int* __restrict__ info, int want_inv, int64_t sA, int64_t sD,
                                                     int64_t sL, const double* __restrict__ Dsrc) {
    __shared__ double Lp[64][65];
    __shared__ double Xs[64][65];
    __shared__ double rinv[64];
    __shared__ double piv[64];
    A += blockIdx.x * sA;
    Dinv += blockIdx.x * sD;
    ldiag += blockIdx.x * sL;
    info += blockIdx.x * ((sA != 0) ? 1 : 0);
    // Dsrc (optional): the FP64 diagonal tiles (64x64 column-major, one per block
    // row, same batch stride as Dinv) to factor in place of the stored tile of A
    if (Dsrc) Dsrc += blockIdx.x * sD + k * 64 * 64;
    diag_block64<T, false>(A, ld, k, Dinv, ldiag, info, want_inv, Dsrc, Lp, Xs, rinv, piv);
}

// ---------------------------------------------------------------------------
// One outer panel of the FP64 blocked Cholesky in a single launch: columns
// [K0, K0 + 64nc), nc <= kPanelMaxBlocks column blocks, all rows from K0.  Workgroup rb owns
// the 64-row block rb (rows K0 + 64rb ..) for the whole panel; for each column
// block c it (1) waits for diagonal block c, (2) solves its block against the
// block's inverse, L_{rb,c} = A_{rb,c} Dinv_c^T, and (3) subtracts
// L_{rb,c} L_{c2,c}^T from its blocks in the later panel columns c2 <= rb,
// reading L_{c2,c} once workgroup c2 has published it.  Workgroup c factors
// diagonal block c (diag_block64) when its own updates are done.  The chain
// per column block is then one diagonal factor, one 64x64 solve and one 64x64
// update inside the workgroups, without the kernel boundaries of the separate
// factor / panel-solve / panel-update launches.
//
// rhs (optional, FP64 storage only): a vector b of N rows that leaves as
// y = L^-1 b, the forward solve carried along: workgroup c stores
// y_c = Dinv_c b_c with diagonal block c, and every workgroup below subtracts
// L_{rb,c} y_c from its block of b after its block solve.  Only the backward
// solve is then left to the caller.
//
// Ordering: each workgroup takes a ticket (atomic counter) as its logical row
// block, so block rb is held by a workgroup that started after those of every
// lower block; all waits are on lower blocks, which have started and do not
// wait on higher ones, so every wait ends.  Flags carry the launch's epoch
// (release after the block's stores, acquire before its loads, agent scope).
// A wait that outlasts maxpoll polls (2^20 by default: a second or more, far
// beyond any legitimate wait) stores kPanelWaitTimeout in info and leaves the
// kernel; the callers report it as a device error.
//
// sync (kPanelSyncWords words): [0, P) diagonal-block flags, [P, P + P^2)
// flags of L_{c2,c} (P + P c2 + c), P = kPanelMaxBlocks, and the ticket counter
// in the last word; base: tickets handed out before this launch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool panel_wait(unsigned* f, unsigned epoch, int* info, int* s_ok, unsigned maxpoll) {
    if (threadIdx.x == 0) {
        int ok = 1;
        unsigned n = 0;
        // poll with plain agent-scope loads, then one acquire fence (its cache
        // invalidation once per wait rather than once per poll)
        while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
            if (n++ >= maxpoll) {
                ok = 0;
                atomicCAS(info, 0, kPanelWaitTimeout);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

// Every wave makes its own stores visible at agent scope (the workgroup barrier
// alone does not wait for other waves' stores), then one thread sets the flag.
__device__ __forceinline__ void panel_publish(unsigned* f, unsigned epoch) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(f, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// S[k][r] = M(r, k) for a 64x64 block of a column-major matrix (leading dimension ld)
template <typename T>
__device__ __forceinline__ void blk_load(double (*S)[65], const T* M, int64_t ld) {
    const int r = threadIdx.x & 63, kq = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 16; ++j) S[kq + 4 * j][r] = (double)M[r + (int64_t)(kq + 4 * j) * ld];
}

// the same block in two steps: its values into registers (a load that can be
// issued early), then into S
template <typename T>
__device__ __forceinline__ void blk_fetch(T (&v)[16], const T* M, int64_t ld) {
    const int r = threadIdx.x & 63, kq = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = M[r + (int64_t)(kq + 4 * j) * ld];
}

template <typename T>
__device__ __forceinline__ void blk_put(double (*S)[65], const T (&v)[16]) {
    const int r = threadIdx.x & 63, kq = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 16; ++j) S[kq + 4 * j][r] = (double)v[j];
}

// M(r, k) = S[k][r] rounded to T; S keeps the rounded value (each thread
// rewrites only the elements it stores)
template <typename T>
__device__ __forceinline__ void blk_store(T* M, int64_t ld, double (*S)[65]) {
    const int r = threadIdx.x & 63, kq = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const T v = (T)S[kq + 4 * j][r];
        M[r + (int64_t)(kq + 4 * j) * ld] = v;
        S[kq + 4 * j][r] = (double)v;
    }
}

// M(r, k) = M(r, k) - S[k][r], formed in FP64 and rounded once to T.  All
// loads are issued before the first store: interleaved, each load would wait
// for the previous store (the compiler cannot rule out that two columns alias
// for an arbitrary ld), sixteen global round trips in a row.
template <typename T>
__device__ __forceinline__ void blk_sub(T* M, int64_t ld, double (*S)[65]) {
    const int r = threadIdx.x & 63, kq = threadIdx.x >> 6;
    T v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = M[r + (int64_t)(kq + 4 * j) * ld];
#pragma unroll
    for (int j = 0; j < 16; ++j) M[r + (int64_t)(kq + 4 * j) * ld] = (T)((double)v[j] - S[kq + 4 * j][r]);
}

// C(r, j) = sum_k P(r, k) Q(j, k) with P, Q in LDS as [k][row]; wave w computes
// the 32x32 quadrant (w / 2, w % 2) as 2 x 2 tiles of v_mfma_f64_16x16x4, then
// (after a barrier) writes C to S as S[j][r] (SUB: subtracts it from S).
// QG: Q is read from global memory instead, Qg (64x64 column-major, Q(j, k)
// at Qg[j + 64k]).
template <bool QG = false, bool SUB = false>
__device__ __forceinline__ void blk_mm_nt(double (*P)[65], double (*Q)[65], double (*S)[65],
                                          const double* __restrict__ Qg = nullptr) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, g = lane >> 4, r0 = 32 * (w >> 1), j0 = 32 * (w & 1);
    d4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int k0 = 0; k0 < 64; k0 += 4) {
        double a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = P[k0 + g][r0 + 16 * i + li];
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = QG ? Qg[(j0 + 16 * j + li) + 64 * (k0 + g)] : Q[k0 + g][j0 + 16 * j + li];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();  // S may be P or Q
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double& d = S[j0 + 16 * j + li][r0 + 16 * i + g + 4 * v];
                d = SUB ? d - acc[i][j][v] : acc[i][j][v];
            }
    __syncthreads();
}

// b_rb -= L_{rb,c} y_c for the 64-row blocks bc (y_c) and brb, L_{rb,c} in S as S[k][r]
__device__ __forceinline__ void rhs_update(const double* bc, double* brb, double (*S)[65], double* vb,
                                           double* red) {
    if (threadIdx.x < 64) vb[threadIdx.x] = bc[threadIdx.x];
    __syncthreads();
    const int e = threadIdx.x & 63, pq = threadIdx.x >> 6;
    double sum = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) sum += S[16 * pq + u][e] * vb[16 * pq + u];
    red[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x < 64) brb[e] -= red[e] + red[e + 64] + red[e + 128] + red[e + 192];
    __syncthreads();
}

//
// T = float (the mixed-precision form): A is FP32; Dacc holds FP64 copies of
// the diagonal tiles (64x64 column-major, tile j at Dacc + 4096j), which the
// diagonal factor reads instead of A's tiles and which receive the updates of
// the diagonal blocks.  Block products are FP64 MFMA on the stored FP32 values
// (exact products, FP64 sums) with one rounding to FP32 per stored result.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_potrf_panel(
    T* __restrict__ A, int64_t ld, int64_t K0, int nc,
                                                     double* __restrict__ Dinv, double* __restrict__ ldiag,
                                                     int* __restrict__ info, unsigned* __restrict__ sync,
                                                     unsigned epoch, unsigned base, double* __restrict__ Dacc,
                                                     unsigned maxpoll, double* __restrict__ rhs) {
    __shared__ double Ps[64][65];
    __shared__ double Qs[64][65];
    __shared__ double rinv[64];
    __shared__ double piv[64];
    __shared__ double red[256];
    __shared__ double vb[64];
    __shared__ unsigned s_id;
    __shared__ int s_ok;
    __shared__ int s_col;
    if (threadIdx.x == 0) s_id = atomicAdd(&sync[kPanelSyncWords - 1], 1u) - base;
    __syncthreads();
    const int rb = (int)s_id;
    const int64_t kb0 = K0 / 64;
    T* Arow = A + K0 + 64 * (int64_t)rb;  // row block rb, column 0
    // kNext: the owner of the next diagonal block keeps that block's last
    // update in LDS (below).  FP64 storage only: with FP32 storage the
    // registers it takes cost more than it saves (mixed-precision fits
    // measured 10 % slower).
    constexpr bool kNext = sizeof(T) == 8;
    for (int c = 0; c < nc && c < rb; ++c) {
        const int64_t kc = kb0 + c;
        T* Ab = Arow + (K0 + 64 * c) * ld;  // A_{rb,c}
        if (kNext && rb == c + 1 && rb < nc) {
            // This workgroup factors diagonal block c + 1 next, and its only update
            // left is that block's own.  A_{rb,c} and the diagonal block are
            // final before the wait (their updates are this workgroup's own
            // stores, made before the barriers since) and are staged into Ps
            // and Qs while it waits; the solve reads Dinv_c straight from
            // global memory, the update is subtracted in LDS, and the factor
            // starts from Qs (FP64).  L_{rb,c} is published after the update
            // (its stores have landed by then).
            blk_load(Ps, Ab, ld);
            blk_load(Qs, A + (kb0 + rb) * 64 * (ld + 1), ld);
            if (!panel_wait(&sync[c], epoch, info, &s_ok, maxpoll)) return;
            blk_mm_nt<true, false>(Ps, Ps, Ps, Dinv + kc * 4096);  // Ps[j][r] = L_{rb,c}(r, j)
            blk_store(Ab, ld, Ps);
            __syncthreads();
            if (rhs) rhs_update(rhs + K0 + 64 * c, rhs + K0 + 64 * rb, Ps, vb, red);
            blk_mm_nt<false, true>(Ps, Ps, Qs);  // Qs -= L_{rb,c} L_{rb,c}^T
            panel_publish(&sync[kPanelMaxBlocks + kPanelMaxBlocks * rb + c], epoch);
            continue;
        }
        // A_{rb,c} is final before the wait: its updates are this workgroup's own
        // stores, made by the same threads (the blk_sub layout)
        T pre[16];
        blk_fetch(pre, Ab, ld);
        if (!panel_wait(&sync[c], epoch, info, &s_ok, maxpoll)) return;
        blk_load(Qs, Dinv + kc * 4096, 64);
        blk_put(Ps, pre);
        __syncthreads();
        blk_mm_nt(Ps, Qs, Ps);  // Ps[j][r] = L_{rb,c}(r, j)
        blk_store(Ab, ld, Ps);  // (Ps now holds the stored values)
        __syncthreads();
        if (rhs) rhs_update(rhs + K0 + 64 * c, rhs + K0 + 64 * rb, Ps, vb, red);
        if (rb < nc) panel_publish(&sync[kPanelMaxBlocks + kPanelMaxBlocks * rb + c], epoch);
        for (int c2 = c + 1; c2 < nc && c2 <= rb; ++c2) {
            if (c2 == rb) {
                blk_mm_nt(Ps, Ps, Qs);  // the block's own diagonal block
            } else {
                if (!panel_wait(&sync[kPanelMaxBlocks + kPanelMaxBlocks * c2 + c], epoch, info, &s_ok, maxpoll)) return;
                blk_load(Qs, A + K0 + 64 * c2 + (K0 + 64 * c) * ld, ld);  // L_{c2,c}
                __syncthreads();
                blk_mm_nt(Ps, Qs, Qs);
            }
            if (Dacc && c2 == rb)
                blk_sub(Dacc + (kb0 + c2) * 4096, 64, Qs);  // the FP64 copy the diagonal factor reads
            else
                blk_sub(Arow + (K0 + 64 * c2) * ld, ld, Qs);
            __syncthreads();
        }
    }
    if (rb < nc) {
        // diagonal block rb with all its updates; with kNext in Qs: left there
        // by the last column (rb > 0), or the stored block (no updates inside
        // the panel)
        const int c = rb;
        const int64_t kc = kb0 + c;
        if (kNext && rb == 0) {
            blk_load(Qs, A + kc * 64 * (ld + 1), ld);
            __syncthreads();
        }
        diag_block64<T, true, kNext>(A, ld, kc, Dinv, ldiag, info, 1, Dacc ? Dacc + kc * 4096 : nullptr, Ps, Qs, rinv,
                                     piv, &s_col);
        if (rhs) {  // y_c = L_cc^-1 b_c from the inverse left in Qs (Qs[k][e] = X(e, k))
            double* bc = rhs + K0 + 64 * c;
            if (threadIdx.x < 64) vb[threadIdx.x] = bc[threadIdx.x];
            __syncthreads();
            const int e = threadIdx.x & 63, pq = threadIdx.x >> 6;
            double sum = 0.0;
#pragma unroll
            for (int u = 0; u < 16; ++u) sum += Qs[16 * pq + u][e] * vb[16 * pq + u];
            red[threadIdx.x] = sum;
            __syncthreads();
            if (threadIdx.x < 64) bc[threadIdx.x] = red[e] + red[e + 64] + red[e + 128] + red[e + 192];
        }
        panel_publish(&sync[c], epoch);
    }
}

void potrf_panel_f64(hipStream_t s, double* A, int64_t ld, int64_t K0, int nc, int64_t nrb, double* Dinv,
                     double* ldiag, int* info, unsigned* sync, unsigned epoch, unsigned base, unsigned maxpoll,
                     double* rhs) {
    hipLaunchKernelGGL(k_potrf_panel<double>, dim3((unsigned)nrb), dim3(256), 0, s, A, ld, K0, nc, Dinv, ldiag, info,
                       sync, epoch, base, (double*)nullptr, maxpoll, rhs);
}

void potrf_panel_mixed(hipStream_t s, float* A, int64_t ld, int64_t K0, int nc, int64_t nrb, double* Dinv,
                       double* ldiag, int* info, unsigned* sync, unsigned epoch, unsigned base, double* Dacc,
                       unsigned maxpoll) {
    hipLaunchKernelGGL(k_potrf_panel<float>, dim3((unsigned)nrb), dim3(256), 0, s, A, ld, K0, nc, Dinv, ldiag, info,
                       sync, epoch, base, Dacc, maxpoll, (double*)nullptr);
}

void potrf_diag3_f64(hipStream_t s, double* A, int64_t ld, int64_t k, double* Dinv, double* ldiag, int* info,
                     int want_inv) {
    hipLaunchKernelGGL(k_potrf_diag3<double>, dim3(1), dim3(256), 0, s, A, ld, k, Dinv, ldiag, info, want_inv,
                       (int64_t)0, (int64_t)0, (int64_t)0, (const double*)nullptr);
}

void potrf_diag3_f32(hipStream_t s, float* A, int64_t ld, int64_t k, double* Dinv, double* ldiag, int* info,
                     int want_inv, const double* Dacc) {
    hipLaunchKernelGGL(k_potrf_diag3<float>, dim3(1), dim3(256), 0, s, A, ld, k, Dinv, ldiag, info, want_inv,
                       (int64_t)0, (int64_t)0, (int64_t)0, Dacc);
}

// batch of problems, each with its own info slot (info[b])
void potrf_diag3_f32_batched(hipStream_t s, float* A, int64_t ld, int64_t k, double* Dinv, double* ldiag, int* info,
                             int batch, int64_t sA, int64_t sD, int64_t sL, const double* Dacc) {
    hipLaunchKernelGGL(k_potrf_diag3<float>, dim3((unsigned)batch), dim3(256), 0, s, A, ld, k, Dinv, ldiag, info, 1,
                       sA, sD, sL, Dacc);
}

// ---------------------------------------------------------------------------
// FP64 diagonal tiles of the mixed-precision factorization.  Every FP32 update
// A -= P P^T that reaches diagonal tiles is also applied to their FP64 copies
// D_j -= P_j P_j^T, where P_j is the FP32 panel's rows of block j: the products
// of FP32 values are exact in FP64 and are accumulated in FP64, so the diagonal
// blocks that are factored carry no FP32 accumulation error.
// D: first tile to update (64x64 column-major tiles, consecutive), P: the panel
// rows of that tile (leading dimension ldp), kdepth: panel width (multiple of 64).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dacc_update(double* __restrict__ D, const float* __restrict__ P, int64_t ldp,
                                                     int64_t kdepth, int64_t sD, int64_t sP) {
    // D_j -= P_j P_j^T on v_mfma_f64_16x16x4f64: each wave owns a 32x32 quadrant
    // (2x2 blocks of 16x16); P_j is staged 16 k at a time in LDS as FP64.
    // f64 16x16x4 layouts: A lane l -> A[l%16][k=l/16], B lane l -> B[k=l/16][l%16],
    // D item v of lane l -> row l/16 + 4v, column l%16.
    D += blockIdx.y * sD + (int64_t)blockIdx.x * 64 * 64;
    P += blockIdx.y * sP + (int64_t)blockIdx.x * 64;
    __shared__ double Ps[16][64 + 2];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wr = w >> 1, wc = w & 1;
    const int l16 = lane & 15, lk = lane >> 4;
    typedef double d4 __attribute__((ext_vector_type(4)));
    d4 acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = d4{0.0, 0.0, 0.0, 0.0};
    for (int64_t k0 = 0; k0 < kdepth; k0 += 16) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int kk = (t >> 6) + 4 * q;
            Ps[kk][lane] = (double)P[lane + (k0 + kk) * ldp];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 16; ks += 4) {
            double a[2], b[2];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi) a[bi] = Ps[ks + lk][32 * wr + 16 * bi + l16];
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) b[bj] = Ps[ks + lk][32 * wc + 16 * bj + l16];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
                    acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b[bj], acc[bi][bj], 0, 0, 0);
        }
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 32 * wr + 16 * bi + lk + 4 * v, col = 32 * wc + 16 * bj + l16;
                D[row + col * 64] -= acc[bi][bj][v];
            }
}

void dacc_update(hipStream_t s, double* D, const float* P, int64_t ldp, int64_t ntiles, int64_t kdepth, int batch,
                 int64_t sD, int64_t sP) {
    if (ntiles <= 0 || batch <= 0) return;
    hipLaunchKernelGGL(k_dacc_update, dim3((unsigned)ntiles, (unsigned)batch), dim3(256), 0, s, D, P, ldp, kdepth, sD,
                       sP);
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
// The same solves in one launch: workgroup k (ticket order) owns block i = k
// (forward) or T - 1 - k (transposed).  It subtracts the products with the
// blocks it depends on as each is published (the tile values are loaded
// before the wait), forms x_i = Dinv_i b_i (Dinv_i^T b_i) and publishes x_i.
// Every wait is on a block of a lower ticket, whose workgroup has started and
// waits only on lower tickets, so every wait ends; the summation order is
// fixed (deterministic).  sync: flags [0, T), the ticket counter at
// sync[kSolveMaxBlocks].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_trsv_persist(const double* __restrict__ L, int64_t ld,
                                                      const double* __restrict__ Dinv, int64_t T,
                                                      const double* __restrict__ b, double* __restrict__ x, int trans,
                                                      unsigned* __restrict__ sync, unsigned epoch, unsigned base,
                                                      int* __restrict__ info, unsigned maxpoll) {
    __shared__ double red[256];
    __shared__ double sv[64];
    __shared__ unsigned s_id;
    __shared__ int s_ok;
    const int t = threadIdx.x, e = t & 63, p = t >> 6;
    if (t == 0) s_id = atomicAdd(&sync[kSolveMaxBlocks], 1u) - base;
    __syncthreads();
    const int64_t k = s_id;
    const int64_t i = trans ? T - 1 - k : k;
    // Dinv_i and b_i do not depend on the waits: loaded first, off the chain
    const double* D = Dinv + i * 64 * 64;
    double dv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int c = 16 * p + u;
        dv[u] = trans ? D[c + e * 64] : D[e + c * 64];
    }
    const double bi = (t < 64) ? b[64 * i + t] : 0.0;
    double acc = 0.0;
    for (int64_t n = 0; n < k; ++n) {
        const int64_t j = trans ? T - 1 - n : n;  // the blocks solved before i, in their order
        double lt[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int r = 16 * p + u;
            // forward: L(64i + e, 64j + r); transposed: L(64j + r, 64i + e)
            lt[u] = trans ? L[(64 * j + r) + (64 * i + e) * ld] : L[(64 * i + e) + (64 * j + r) * ld];
        }
        if (!panel_wait(&sync[j], epoch, info, &s_ok, maxpoll)) return;
#pragma unroll
        for (int u = 0; u < 16; ++u) acc += lt[u] * x[64 * j + 16 * p + u];
    }
    red[t] = acc;
    __syncthreads();
    if (t < 64) sv[t] = bi - (red[t] + red[t + 64] + red[t + 128] + red[t + 192]);
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) s += dv[u] * sv[16 * p + u];
    red[t] = s;
    __syncthreads();
    if (t < 64) x[64 * i + t] = red[t] + red[t + 64] + red[t + 128] + red[t + 192];
    panel_publish(&sync[i], epoch);
}

void trsv_persist_f64(hipStream_t s, const double* L, int64_t ld, const double* Dinv, int64_t N, const double* b,
                      double* x, int trans, unsigned* sync, unsigned epoch, unsigned base, int* info,
                      unsigned maxpoll) {
    const int64_t T = N / 64;
    if (T <= 0) return;
    hipLaunchKernelGGL(k_trsv_persist, dim3((unsigned)T), dim3(256), 0, s, L, ld, Dinv, T, b, x, trans, sync, epoch,
                       base, info, maxpoll);
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

// y = A v for a symmetric FP64 matrix held as its lower 64x64 tiles (column-
// major, leading dimension ld; diagonal tiles complete), reading every stored
// tile once.  The lower tiles are grouped into G x G super-tiles; block
// (RI, CJ) of the lower super-tile grid (RI >= CJ) takes tiles (ti, tj) of its
// super-tile with tj <= ti.  An off-diagonal tile serves both triangles: A_ij v_j
// goes to output block i (the "row" role) and A_ij^T v_i to output block j (the
// "column" role); a diagonal tile serves the row role alone.  Thread (r, p)
// reads row r of columns 16p..16p+15 of each tile, so every load is a 512-byte
// column segment per wave.  Row-role sums are kept per row tile and reduced over
// the four waves at the end; column-role sums are kept for one column tile at a
// time (16 per thread) and reduced over the 64 rows by a butterfly that halves
// the values held per lane at each exchange.  Both go to partial slots of their
// own: part[CJ][i] (row role of column group CJ) and part[S + RI][j] (column role
// of row group RI), S = number of super-tile rows.  A second pass adds the slots
// that exist for each output block in a fixed order, so the result does not
// depend on scheduling.
static int symv_group(int batch) { return batch >= 4 ? 4 : 2; }

int symv_slots(int64_t N, int batch) {
    const int64_t T = N / 64, G = symv_group(batch), S = (T + G - 1) / G;
    return (int)(2 * (S > 0 ? S : 1));
}

// Sum of v over the 64 lanes for 16 values per lane: after the exchanges with
// lanes l^32, l^16, l^8, l^4 a lane holds one value, the partial of value
// u(l) = 8 b5 + 4 b4 + 2 b3 + b2 (bk = bit k of l) over the 16 lanes that agree
// with it in bits 2..5; the exchanges with l^2 and l^1 complete the sum.
__device__ __forceinline__ double butterfly16(double (&v)[16], int lane) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const bool hi = (lane & 32) != 0;
        const double keep = hi ? v[u + 8] : v[u], give = hi ? v[u] : v[u + 8];
        v[u] = keep + __shfl_xor(give, 32, 64);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const bool hi = (lane & 16) != 0;
        const double keep = hi ? v[u + 4] : v[u], give = hi ? v[u] : v[u + 4];
        v[u] = keep + __shfl_xor(give, 16, 64);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const bool hi = (lane & 8) != 0;
        const double keep = hi ? v[u + 2] : v[u], give = hi ? v[u] : v[u + 2];
        v[u] = keep + __shfl_xor(give, 8, 64);
    }
    {
        const bool hi = (lane & 4) != 0;
        const double keep = hi ? v[1] : v[0], give = hi ? v[0] : v[1];
        v[0] = keep + __shfl_xor(give, 4, 64);
    }
    double x = v[0];
    x += __shfl_xor(x, 2, 64);
    x += __shfl_xor(x, 1, 64);
    return x;
}

template <int G>
__global__ __launch_bounds__(256) void k_symv2(const double* __restrict__ A, int64_t ld, int64_t T,
                                               const double* __restrict__ v, double* __restrict__ part, int64_t sA,
                                               int64_t sV) {
    const int64_t b = blockIdx.y, S = (T + G - 1) / G;
    A += b * sA;
    v += b * sV;
    part += b * 2 * S * T * 64;
    int64_t RI, CJ;
    lower_tile(blockIdx.x, RI, CJ);
    const int64_t ti0 = RI * G, tj0 = CJ * G;
    const int ni = (int)((ti0 + G <= T) ? G : T - ti0), nj = (int)((tj0 + G <= T) ? G : T - tj0);
    const int t = threadIdx.x, r = t & 63, p = t >> 6;
    __shared__ double vc[G][64];  // v over the column tiles
    __shared__ double red[4][64];
    if (t < 64) {
#pragma unroll
        for (int q = 0; q < G; ++q) vc[q][t] = (q < nj) ? v[(tj0 + q) * 64 + t] : 0.0;
    }
    double vr[G], rowacc[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
        vr[q] = (q < ni) ? v[(ti0 + q) * 64 + r] : 0.0;
        rowacc[q] = 0.0;
    }
    __syncthreads();
    // the column role's output: lane l of wave p ends with column 16p + u(l)
    const int uo = ((r >> 5) & 1) * 8 + ((r >> 4) & 1) * 4 + ((r >> 3) & 1) * 2 + ((r >> 2) & 1);
    double* pcol = part + (S + RI) * T * 64;
#pragma unroll
    for (int qj = 0; qj < G; ++qj) {
        if (qj >= nj) break;  // block-uniform
        const int64_t tj = tj0 + qj;
        double colacc[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) colacc[u] = 0.0;
#pragma unroll
        for (int qi = 0; qi < G; ++qi) {
            const int64_t ti = ti0 + qi;
            if (qi >= ni || tj > ti) continue;  // block-uniform
            const double* At = A + ti * 64 + r + (tj * 64 + 16 * p) * ld;
            double a[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) a[u] = At[(int64_t)u * ld];
            double s = 0.0;
#pragma unroll
            for (int u = 0; u < 16; ++u) s += a[u] * vc[qj][16 * p + u];
            rowacc[qi] += s;
            if (tj < ti) {
#pragma unroll
                for (int u = 0; u < 16; ++u) colacc[u] += a[u] * vr[qi];
            }
        }
        const double cs = butterfly16(colacc, r);
        if ((r & 3) == 0) pcol[tj * 64 + 16 * p + uo] = cs;
    }
    // row role: sum over the four waves, per row tile
    double* prow = part + CJ * T * 64;
#pragma unroll
    for (int qi = 0; qi < G; ++qi) {
        red[p][r] = rowacc[qi];
        __syncthreads();
        if (t < 64 && qi < ni) prow[(ti0 + qi) * 64 + t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
        __syncthreads();
    }
}

// y[i] = sum of the row-role slots CJ <= I and the column-role slots RI >= I of
// output block i (I = its super-tile row), in that order
__global__ void k_symv2_sum(const double* __restrict__ part, int64_t T, int G, double* __restrict__ y, int64_t sV) {
    const int64_t b = blockIdx.y, S = (T + G - 1) / G, N = T * 64;
    part += b * 2 * S * N;
    y += b * sV;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t I = (i / 64) / G;
    double s = 0.0;
    for (int64_t cg = 0; cg <= I; ++cg) s += part[cg * N + i];
    for (int64_t rg = I; rg < S; ++rg) s += part[(S + rg) * N + i];
    y[i] = s;
}

__global__ void k_chunk_sum(const double* __restrict__ part, int64_t N, int nchunk, double* __restrict__ y,
                            int64_t sV) {
    const int64_t b = blockIdx.y;
    part += b * (int64_t)nchunk * N;
    y += b * sV;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * N + i];
    y[i] = s;
}

// part: batch * symv_slots(N, batch) * N doubles (nchunk is not used)
void chunk_sum_f64(hipStream_t s, const double* part, int64_t N, int nchunk, double* y, int batch, int64_t sV);
void symv_lt_f64(hipStream_t s, const double* A, int64_t N, const double* v, double* y, double* part, int nchunk,
                 int batch, int64_t sA, int64_t sV) {
    (void)nchunk;
    const int64_t T = N / 64;
    const int G = symv_group(batch);
    const int64_t S = (T + G - 1) / G;
    const dim3 grid((unsigned)(S * (S + 1) / 2), (unsigned)batch);
    if (G == 4) hipLaunchKernelGGL(k_symv2<4>, grid, dim3(256), 0, s, A, N, T, v, part, sA, sV);
    else hipLaunchKernelGGL(k_symv2<2>, grid, dim3(256), 0, s, A, N, T, v, part, sA, sV);
    hipLaunchKernelGGL(k_symv2_sum, dim3((unsigned)((N + 255) / 256), (unsigned)batch), dim3(256), 0, s, part, T, G, y,
                       sV);
}

// y[b][i] = sum_c part[b][c][i] (fixed order)
void chunk_sum_f64(hipStream_t s, const double* part, int64_t N, int nchunk, double* y, int batch, int64_t sV) {
    hipLaunchKernelGGL(k_chunk_sum, dim3((unsigned)((N + 255) / 256), (unsigned)batch), dim3(256), 0, s, part, N,
                       nchunk, y, sV);
}

// The weight matrix of the latent-likelihood gradient, in place over the lower
// tiles of B^-1 (every element of a lower tile, diagonal tiles complete):
// W_ik = s_i B^-1_ik s_k - (t_i u_k + u_i t_k) for i, k < n (t, u may be null:
// no rank-two term), zero where i or k is padding.  With s = V^-1/2,
// s B^-1 s = (K + V)^-1.
__global__ __launch_bounds__(256) void k_form_wgrad(double* __restrict__ B, int64_t ld, int64_t n,
                                                    const double* __restrict__ sv, const double* __restrict__ t,
                                                    const double* __restrict__ u) {
    int64_t ti, tj;
    lower_tile(blockIdx.x, ti, tj);
    const int e = threadIdx.x & 63, p = threadIdx.x >> 6;
    const int64_t i = ti * 64 + e;
    const double si = (i < n) ? sv[i] : 0.0;
    const double tii = (t && i < n) ? t[i] : 0.0, ui = (u && i < n) ? u[i] : 0.0;
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
        const int64_t k = tj * 64 + 16 * p + q;
        double* b = B + i + k * ld;
        if (i < n && k < n) {
            double w = si * (*b) * sv[k];
            if (t) w -= tii * u[k] + ui * t[k];
            *b = w;
        } else {
            *b = 0.0;
        }
    }
}

void form_wgrad_f64(hipStream_t s, double* B, int64_t N, int64_t n, const double* sv, const double* t,
                    const double* u) {
    const int64_t T = N / 64;
    hipLaunchKernelGGL(k_form_wgrad, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, B, N, n, sv, t, u);
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
// A = I over all T x T tiles: the triangular GEMM modes read whole tiles, so the
// upper tiles of a triangular factor must hold zeros.
__global__ void k_set_identity(double* __restrict__ A, int64_t ld, int64_t T) {
    const int64_t ti = blockIdx.x % T, tj = blockIdx.x / T;
    const int t = threadIdx.x, i = t & 63;
    for (int c = t >> 6; c < 64; c += 4) {
        const int64_t gi = ti * 64 + i, gj = tj * 64 + c;
        A[gi + gj * ld] = (gi == gj) ? 1.0 : 0.0;
    }
}

void set_identity_f64(hipStream_t s, double* A, int64_t ld, int64_t N) {
    const int64_t T = N / 64;
    hipLaunchKernelGGL(k_set_identity, dim3((unsigned)(T * T)), dim3(256), 0, s, A, ld, T);
}

// Column j = c0 + blockIdx.x of X (N x N, ld N): zero above its 64x64
// diagonal block, the block's column from Dinv (64x64 column-major per block,
// zero above its diagonal) on it; the rows below are left to the caller.
__global__ __launch_bounds__(256) void k_trtri_leaves(double* __restrict__ X, int64_t N,
                                                      const double* __restrict__ Dinv, int64_t c0) {
    const int64_t j = c0 + blockIdx.x, kb = j / 64, r0 = kb * 64;
    const double* D = Dinv + kb * 4096 + (j - r0) * 64;
    for (int64_t i = threadIdx.x; i < r0 + 64; i += 256) X[i + j * N] = (i < r0) ? 0.0 : D[i - r0];
}

void trtri_leaves_f64(hipStream_t s, double* X, int64_t N, const double* Dinv, int64_t c0, int64_t ncols) {
    if (ncols <= 0) return;
    hipLaunchKernelGGL(k_trtri_leaves, dim3((unsigned)ncols), dim3(256), 0, s, X, N, Dinv, c0);
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

// B = I + S K S over lower tiles (S = diag(sv), identity on padding).  Optional
// outputs for the mixed-precision form: B32 = fl32(B) and dacc = the unrounded
// FP64 diagonal tiles (64x64 column-major, one per block row).
__global__ void k_form_B(const double* __restrict__ K, double* __restrict__ B, int64_t ld, int64_t n,
                         const double* __restrict__ sv, float* __restrict__ B32, double* __restrict__ dacc) {
    int64_t ti, tj;
    lower_tile(blockIdx.x, ti, tj);
    const int t = threadIdx.x, i = t & 63;
    const int64_t gi = ti * 64 + i;
    const double si = (gi < n) ? sv[gi] : 0.0;
    double* dt = (dacc && ti == tj) ? dacc + ti * 64 * 64 : nullptr;
    for (int c = t >> 6; c < 64; c += 4) {
        const int64_t gj = tj * 64 + c;
        double v;
        if (gi < n && gj < n) v = si * K[gi + gj * ld] * sv[gj] + (gi == gj ? 1.0 : 0.0);
        else v = (gi == gj) ? 1.0 : 0.0;
        B[gi + gj * ld] = v;
        if (B32) B32[gi + gj * ld] = (float)v;
        if (dt) dt[i + c * 64] = v;
    }
}

void form_B_f64(hipStream_t s, const double* K, double* B, int64_t ld, int64_t n, int64_t N, const double* sv) {
    const int64_t T = N / 64;
    hipLaunchKernelGGL(k_form_B, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, K, B, ld, n, sv, (float*)nullptr,
                       (double*)nullptr);
}

void form_B_mixed(hipStream_t s, const double* K, double* B64, float* B32, double* dacc, int64_t ld, int64_t n,
                  int64_t N, const double* sv) {
    const int64_t T = N / 64;
    hipLaunchKernelGGL(k_form_B, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, K, B64, ld, n, sv, B32, dacc);
}

// ---------------------------------------------------------------------------
// Restore a saved factor (gp_load).  In place over L (N x N, column-major): the
// n x n part holds C_chol = diag(sqrt V) chol(B)^T as saved; it becomes
// L = diag(1/sqrt V) C_chol with zeros above the diagonal, the padding becomes
// the identity, and sv[i] = 1/sqrt(V_i).  One workgroup per column.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_restore_factor(double* __restrict__ L, int64_t n, int64_t N,
                                                        const double* __restrict__ V, double* __restrict__ sv) {
    const int64_t j = blockIdx.x;
    for (int64_t i = threadIdx.x; i < N; i += 256) {
        double v;
        if (i < n && j < n) v = (i >= j) ? L[i + j * N] / sqrt(V[i]) : 0.0;
        else v = (i == j) ? 1.0 : 0.0;
        L[i + j * N] = v;
        if (j == 0 && i < n) sv[i] = 1.0 / sqrt(V[i]);
    }
}

void restore_factor_f64(hipStream_t s, double* L, int64_t n, int64_t N, const double* V, double* sv) {
    hipLaunchKernelGGL(k_restore_factor, dim3((unsigned)N), dim3(256), 0, s, L, n, N, V, sv);
}

// Inverses of the 64x64 diagonal blocks of an existing lower-triangular factor
// (the part of k_potrf_diag3 that follows the factorisation): Dinv + k*4096
// receives L_kk^-1 (column-major, ld 64, zeros above the diagonal) and
// ldiag[64k + i] = log L_ii.  Lane t forward-substitutes column t.  A pivot
// that is not positive (or NaN) sets info to its 1-based index if info is 0.
__global__ __launch_bounds__(64) void k_trtri_diag64(const double* __restrict__ L, int64_t ld,
                                                     double* __restrict__ Dinv, double* __restrict__ ldiag,
                                                     int* __restrict__ info) {
    __shared__ double A[64][65];  // A[col][row] = L_kk(row, col)
    __shared__ double X[64][65];  // X[col][row] = L_kk^-1(row, col)
    const int64_t k = blockIdx.x;
    const int t = threadIdx.x;
    const double* blk = L + k * 64 * (ld + 1);
    for (int c = 0; c < 64; ++c) A[c][t] = blk[t + (int64_t)c * ld];
    __syncthreads();
    for (int i = 0; i < 64; ++i) {
        double v = 0.0;
        if (i >= t) {
            double acc = (i == t) ? 1.0 : 0.0;
            for (int q = t; q < i; ++q) acc -= A[q][i] * X[t][q];
            v = acc / A[i][i];
        }
        X[t][i] = v;
    }
    __syncthreads();
    double* D = Dinv + k * 64 * 64;
    for (int c = 0; c < 64; ++c) D[t + c * 64] = X[c][t];
    const double p = A[t][t];
    ldiag[k * 64 + t] = log(p);
    const unsigned long long bad = __ballot(!(p > 0.0));
    if (t == 0 && bad) atomicCAS(info, 0, (int)(k * 64 + __ffsll((long long)bad)));
}

void trtri_diag64_f64(hipStream_t s, const double* L, int64_t ld, int64_t T, double* Dinv, double* ldiag, int* info) {
    hipLaunchKernelGGL(k_trtri_diag64, dim3((unsigned)T), dim3(64), 0, s, L, ld, Dinv, ldiag, info);
}

}  // namespace gpl
