// Covariance builds for gfx950: the device restatement of src/cf.cpp
// (cf_sexp_c :9-42, cf_matern32_c :48-80, cf_matern52_c :86-119, cf_nn_c :125-162)
// and of the R-side composition in R/cf.R:402-497 (sum of terms, products,
// periodic feature map), extended with ARD length-scales and d K / d log(theta).
//
// Points are staged feature-major in LDS (Fs[f*64 + i]) so that a wavefront reads
// one feature of 64 consecutive points per instruction; output tiles are written
// column by column, 64 consecutive rows per store (512 coalesced bytes).
#include "gpl_internal.h"

#include <cmath>

namespace gpl {

__device__ __forceinline__ int64_t xcd_remap_c(int64_t b, int64_t nb) {
    const int64_t per = nb / 8, rem = nb % 8;
    const int64_t x = b % 8, q = b / 8;
    return (x < rem) ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}

__device__ __forceinline__ void lower_tile_c(int64_t t, int64_t& ti, int64_t& tj) {
    int64_t r = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    ti = r;
    tj = t - r * (r + 1) / 2;
}

constexpr double kSqrt3 = 1.7320508075688772935;
constexpr double kSqrt5 = 2.2360679774997896964;
constexpr double kTwoOverPi = 0.63661977236758134308;

// stationary profile k = g(rho), rho = sum_f (x_f - x'_f)^2 / l_f^2
__device__ __forceinline__ double stat_g(int kind, double rho, double m2) {
    if (kind == GPL_CF_SEXP) return m2 * exp(-0.5 * rho);
    const double r = sqrt(rho);
    if (kind == GPL_CF_MATERN32) {
        const double u = kSqrt3 * r;
        return m2 * (1.0 + u) * exp(-u);
    }
    const double u = kSqrt5 * r;
    return m2 * (1.0 + u + 5.0 * rho / 3.0) * exp(-u);
}

// (g, dg/drho)
__device__ __forceinline__ void stat_gd(int kind, double rho, double m2, double& g, double& dg) {
    if (kind == GPL_CF_SEXP) {
        g = m2 * exp(-0.5 * rho);
        dg = -0.5 * g;
        return;
    }
    const double r = sqrt(rho);
    if (kind == GPL_CF_MATERN32) {
        const double u = kSqrt3 * r, e = exp(-u);
        g = m2 * (1.0 + u) * e;
        dg = -1.5 * m2 * e;
        return;
    }
    const double u = kSqrt5 * r, e = exp(-u);
    g = m2 * (1.0 + u + 5.0 * rho / 3.0) * e;
    dg = -(5.0 / 6.0) * m2 * (1.0 + u) * e;
}

// fi / fj: feature vectors with strides si / sj
__device__ __forceinline__ double term_value(const DevTerm& T, const double* fi, int si, const double* fj,
                                             int sj) {
    switch (T.kind) {
        case GPL_CF_CONST:
            return T.m2;
        case GPL_CF_LIN: {
            double s = 0.0;
            for (int f = 0; f < T.nf; ++f) s += fi[(T.f0 + f) * si] * fj[(T.f0 + f) * sj];
            return T.m2 * s;
        }
        case GPL_CF_NN: {
            double a = 0.0, b = 0.0, c = 0.0;
            for (int f = 0; f < T.nf; ++f) {
                const double u = fi[(T.f0 + f) * si], v = fj[(T.f0 + f) * sj];
                a += u * v; b += u * u; c += v * v;
            }
            a = T.b0 + a * T.s2; b = T.b0 + b * T.s2; c = T.b0 + c * T.s2;
            return T.m2 * kTwoOverPi * asin(2.0 * a / sqrt((1.0 + 2.0 * b) * (1.0 + 2.0 * c)));
        }
        default: {  // SEXP, MATERN32, MATERN52, PERIODIC (base on the sin/cos features)
            double rho = 0.0;
            for (int f = 0; f < T.nf; ++f) {
                const double dd = fi[(T.f0 + f) * si] - fj[(T.f0 + f) * sj];
                rho += dd * dd * T.il2[f];
            }
            return stat_g(T.kind == GPL_CF_PERIODIC ? T.base_kind : T.kind, rho, T.m2);
        }
    }
}

__device__ __forceinline__ double prog_value(const DevProg* __restrict__ P, const double* fi, int si,
                                             const double* fj, int sj) {
    const int nt = P->nterms;
    double total = 0.0, cur = 0.0;
    for (int k = 0; k < nt; ++k) {
        const DevTerm& T = P->t[k];
        const double v = term_value(T, fi, si, fj, sj);
        if (k == 0) cur = v;
        else if (T.op == GPL_OP_PROD) cur *= v;
        else { total += cur; cur = v; }
    }
    return total + cur;
}

// d term / d log(theta_local) of one term
__device__ double term_dparam(const DevTerm& T, int lp, const double* fi, int si, const double* fj, int sj) {
    switch (T.kind) {
        case GPL_CF_CONST:
            return 2.0 * T.m2;
        case GPL_CF_LIN:
            return 2.0 * term_value(T, fi, si, fj, sj);
        case GPL_CF_NN: {
            double a = 0.0, b = 0.0, c = 0.0;
            for (int f = 0; f < T.nf; ++f) {
                const double u = fi[(T.f0 + f) * si], v = fj[(T.f0 + f) * sj];
                a += u * v; b += u * u; c += v * v;
            }
            a = T.b0 + a * T.s2; b = T.b0 + b * T.s2; c = T.b0 + c * T.s2;
            const double D = (1.0 + 2.0 * b) * (1.0 + 2.0 * c);
            const double sD = sqrt(D);
            const double u = 2.0 * a / sD;
            if (lp == 2) return 2.0 * T.m2 * kTwoOverPi * asin(u);
            const double dk = T.m2 * kTwoOverPi / sqrt(1.0 - u * u);
            double da, db, dc;
            if (lp == 0) { da = db = dc = 2.0 * T.b0; }
            else { da = 2.0 * a; db = 2.0 * b; dc = 2.0 * c; }
            return dk * (2.0 * da / sD - u * (db / (1.0 + 2.0 * b) + dc / (1.0 + 2.0 * c)));
        }
        default: {
            const int kind = (T.kind == GPL_CF_PERIODIC) ? T.base_kind : T.kind;
            int l = lp;
            double rho = 0.0, dper = 0.0, dsel = 0.0;
            const bool per = (T.kind == GPL_CF_PERIODIC);
            if (per) l = lp - 1;
            for (int f = 0; f < T.nf; ++f) {
                const double dd = fi[(T.f0 + f) * si] - fj[(T.f0 + f) * sj];
                const double t2 = dd * dd * T.il2[f];
                rho += t2;
                if (T.nl > 1 && f == l) dsel = t2;
                if (per && lp == 0) {
                    const double ddd = fi[(T.df0 + f) * si] - fj[(T.df0 + f) * sj];
                    dper += dd * ddd * T.il2[f];
                }
            }
            double g, dg;
            stat_gd(kind, rho, T.m2, g, dg);
            if (per && lp == 0) return dg * 2.0 * dper;
            if (l == T.nl) return 2.0 * g;                 // magn
            return dg * (-2.0) * ((T.nl > 1) ? dsel : rho);  // lscale
        }
    }
}

// d K / d log(theta_p) for global gradient slot p (product rule over the summand)
__device__ double prog_dparam(const DevProg* __restrict__ P, int p, const double* fi, int si,
                              const double* fj, int sj) {
    const int nt = P->nterms;
    int owner = -1;
    for (int k = 0; k < nt; ++k)
        if (p >= P->t[k].g0 && p < P->t[k].g0 + P->t[k].ng) owner = k;
    if (owner < 0) return 0.0;
    int s0 = owner;
    while (s0 > 0 && P->t[s0].op == GPL_OP_PROD) --s0;
    int s1 = owner + 1;
    while (s1 < nt && P->t[s1].op == GPL_OP_PROD) ++s1;
    double prod = 1.0;
    for (int k = s0; k < s1; ++k)
        if (k != owner) prod *= term_value(P->t[k], fi, si, fj, sj);
    return prod * term_dparam(P->t[owner], p - P->t[owner].g0, fi, si, fj, sj);
}

// ---------------------------------------------------------------------------
// Features: raw columns -> feature columns (periodic: sin/cos and d/dlog p)
// ---------------------------------------------------------------------------
__global__ void k_features(const double* __restrict__ x, int64_t n, int64_t ldx, const DevProg* __restrict__ P,
                           double* __restrict__ F, int64_t ldf, int64_t N, const int64_t* __restrict__ xoff,
                           const int64_t* __restrict__ nvec, int64_t sF) {
    const int64_t b = blockIdx.y;  // batch: problem b has n = nvec[b] rows at x + xoff[b]
    if (nvec) {
        n = nvec[b];
        ldx = n;
        x += xoff[b];
        P += b;
        F += b * sF;
    }
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int nt = P->nterms;
    for (int k = 0; k < nt; ++k) {
        const DevTerm& T = P->t[k];
        if (T.kind == GPL_CF_CONST) continue;
        for (int c = 0; c < T.nc; ++c) {
            const double v = (i < n) ? x[i + (int64_t)(T.c0 + c) * ldx] : 0.0;
            if (T.kind == GPL_CF_PERIODIC) {
                const double a = T.w * v;
                const double sa = sin(a), ca = cos(a);
                const bool ok = i < n;
                F[i + (int64_t)(T.f0 + c) * ldf] = ok ? sa : 0.0;
                F[i + (int64_t)(T.f0 + T.nc + c) * ldf] = ok ? ca : 0.0;
                F[i + (int64_t)(T.df0 + c) * ldf] = ok ? -a * ca : 0.0;
                F[i + (int64_t)(T.df0 + T.nc + c) * ldf] = ok ? a * sa : 0.0;
            } else {
                F[i + (int64_t)(T.f0 + c) * ldf] = v;
            }
        }
    }
}

void launch_features(hipStream_t s, const double* x, int64_t n, int64_t ldx, int ncols, const DevProg* prog_dev,
                     const DevProg& prog, double* F, int64_t ldf) {
    (void)ncols;
    (void)prog;
    const int64_t N = ldf;
    hipLaunchKernelGGL(k_features, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, x, n, ldx, prog_dev, F, ldf,
                       N, (const int64_t*)nullptr, (const int64_t*)nullptr, (int64_t)0);
}

void launch_features_batched(hipStream_t s, const double* x, const int64_t* xoff, const int64_t* nvec,
                             const DevProg* progs, double* F, int64_t N, int64_t sF, int batch) {
    hipLaunchKernelGGL(k_features, dim3((unsigned)((N + 255) / 256), (unsigned)batch), dim3(256), 0, s, x, (int64_t)0,
                       (int64_t)0, progs, F, N, N, xoff, nvec, sF);
}

// ---------------------------------------------------------------------------
// Symmetric build over lower 64x64 tiles
// ---------------------------------------------------------------------------
template <typename OutT>
__global__ __launch_bounds__(256) void k_build_sym(const DevProg* __restrict__ P, const double* __restrict__ F,
                                                   int64_t ldf, int nfeat, int64_t n, double jitter, int mode,
                                                   double sscal, const double* __restrict__ svec,
                                                   OutT* __restrict__ out, int64_t ld, int64_t sF, int64_t sO,
                                                   const int64_t* __restrict__ nvec,
                                                   const double* __restrict__ scal, double* __restrict__ dacc,
                                                   int64_t sDa, double* __restrict__ out64, int64_t sO64) {
    extern __shared__ double sm[];
    double* Fr = sm;               // [nfeat][64]
    double* Fc = sm + nfeat * 64;  // [nfeat][64]
    {
        const int64_t b = blockIdx.y;  // batch
        P += (nvec ? b : 0);
        F += b * sF;
        out += b * sO;
        if (dacc) dacc += b * sDa;
        if (out64) out64 += b * sO64;
        if (nvec) n = nvec[b];
        if (scal) sscal = scal[b];
    }
    int64_t ti, tj;
    lower_tile_c(xcd_remap_c(blockIdx.x, gridDim.x), ti, tj);
    // dacc (optional): unrounded FP64 copies of the diagonal tiles, 64x64 column-major per block row;
    // out64 (optional): an FP64 copy of the lower tiles (with the FP32 out)
    double* dt = (dacc && ti == tj) ? dacc + ti * 64 * 64 : nullptr;
    const int t = threadIdx.x, i = t & 63, cg = t >> 6;
    for (int e = t; e < nfeat * 64; e += 256) {
        const int f = e >> 6, p = e & 63;
        Fr[e] = F[ti * 64 + p + (int64_t)f * ldf];
        Fc[e] = F[tj * 64 + p + (int64_t)f * ldf];
    }
    __syncthreads();
    const int64_t gi = ti * 64 + i;
    const double si = svec ? ((gi < n) ? svec[gi] : 0.0) : sscal;
#pragma unroll 2
    for (int q = 0; q < 16; ++q) {
        const int c = cg + 4 * q;
        const int64_t gj = tj * 64 + c;
        double v;
        if (gi < n && gj < n) {
            v = prog_value(P, Fr + i, 64, Fc + c, 64);
            if (gi == gj) v += jitter;
            if (mode == 1) {
                const double sj = svec ? svec[gj] : sscal;
                v = si * v * sj + (gi == gj ? 1.0 : 0.0);
            }
        } else {
            v = (gi == gj) ? 1.0 : 0.0;
        }
        out[gi + gj * ld] = (OutT)v;
        if (out64) out64[gi + gj * ld] = v;
        if (dt) dt[i + c * 64] = v;
    }
}

static void build_sym_launch(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F, int64_t n,
                             int64_t N, double jitter, int mode, double sscal, const double* svec, double* out,
                             float* out32, int64_t ld, double* dacc = nullptr, double* out64 = nullptr) {
    const int64_t T = N / 64;
    const int nfeat = hp.nfeat > 0 ? hp.nfeat : 1;
    const size_t shm = (size_t)nfeat * 64 * 2 * sizeof(double);
    dim3 grid((unsigned)(T * (T + 1) / 2));
    if (out)
        hipLaunchKernelGGL(k_build_sym<double>, grid, dim3(256), shm, s, prog, F, N, nfeat, n, jitter, mode, sscal,
                           svec, out, ld, (int64_t)0, (int64_t)0, (const int64_t*)nullptr, (const double*)nullptr,
                           dacc, (int64_t)0, (double*)nullptr, (int64_t)0);
    else
        hipLaunchKernelGGL(k_build_sym<float>, grid, dim3(256), shm, s, prog, F, N, nfeat, n, jitter, mode, sscal,
                           svec, out32, ld, (int64_t)0, (int64_t)0, (const int64_t*)nullptr, (const double*)nullptr,
                           dacc, (int64_t)0, out64, (int64_t)0);
}

// batch of problems of one padded size N: B_b = I + s_b (K_b + jitter I) s_b in FP32
void launch_build_B_f32_batched(hipStream_t s, const DevProg* progs, int nfeat, const double* F, int64_t sF,
                                const int64_t* nvec, int64_t N, double jitter, const double* scal, float* out,
                                int64_t sO, int batch, double* dacc, int64_t sDa, double* out64) {
    const int64_t T = N / 64;
    nfeat = nfeat > 0 ? nfeat : 1;
    const size_t shm = (size_t)nfeat * 64 * 2 * sizeof(double);
    hipLaunchKernelGGL(k_build_sym<float>, dim3((unsigned)(T * (T + 1) / 2), (unsigned)batch), dim3(256), shm, s,
                       progs, F, N, nfeat, (int64_t)0, jitter, 1, 1.0, (const double*)nullptr, out, N, sF, sO, nvec,
                       scal, dacc, sDa, out64, sO);
}

void launch_build_sym_f64(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F, int64_t n,
                          int64_t N, double jitter, int mode, double sscal, const double* svec, double* out,
                          int64_t ld) {
    build_sym_launch(s, prog, hp, F, n, N, jitter, mode, sscal, svec, out, nullptr, ld);
}

void launch_build_sym_f32(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F, int64_t n,
                          int64_t N, double jitter, int mode, double sscal, const double* svec, float* out,
                          int64_t ld, double* dacc, double* out64) {
    build_sym_launch(s, prog, hp, F, n, N, jitter, mode, sscal, svec, nullptr, out, ld, dacc, out64);
}

// ---------------------------------------------------------------------------
// First-order log-determinant correction for a factor of the FP32-rounded B.
// The FP32 path factors Bt, whose off-diagonal tiles are fl32(B) (its diagonal
// tiles are kept in FP64), so log|B| = log|Bt| + tr(Bt^-1 (B - Bt))
// + O(|Bt^-1 (B - Bt)|^2).  Per off-diagonal lower tile this kernel reads the
// FP64 copy of B, takes its FP32 rounding error and accumulates
// 2 (B^-1)_ij (B_ij - fl32(B_ij)).  part[b * tiles + tile] receives the tile sum
// (deterministic order; 0 for diagonal tiles).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_round_corr(const double* __restrict__ B64, const float* __restrict__ W,
                                                    int64_t ld, double* __restrict__ part, int64_t sB, int64_t sW) {
    __shared__ double red[256];
    B64 += blockIdx.y * sB;
    W += blockIdx.y * sW;
    part += blockIdx.y * (int64_t)gridDim.x;
    int64_t ti, tj;
    lower_tile_c(blockIdx.x, ti, tj);
    const int t = threadIdx.x, i = t & 63, cg = t >> 6;
    double acc = 0.0;
    if (ti != tj) {
        const int64_t gi = ti * 64 + i;
#pragma unroll 4
        for (int q = 0; q < 16; ++q) {
            const int64_t gj = tj * 64 + cg + 4 * q;
            const double v = B64[gi + gj * ld];
            acc += (double)W[gi + gj * ld] * (v - (double)(float)v);
        }
    }
    red[t] = 2.0 * acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) part[blockIdx.x] = red[0];
}

void launch_round_corr(hipStream_t s, const double* B64, const float* W, int64_t N, double* part, int batch,
                       int64_t sB, int64_t sW) {
    const int64_t T = N / 64;
    hipLaunchKernelGGL(k_round_corr, dim3((unsigned)(T * (T + 1) / 2), (unsigned)batch), dim3(256), 0, s, B64, W, N,
                       part, sB, sW);
}

// ---------------------------------------------------------------------------
// Cross build: out (Na x Nb) = K(Fa, Fb), zero outside the valid block
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_build_cross(const DevProg* __restrict__ P, const double* __restrict__ Fa,
                                                     int64_t ldfa, int64_t na, const double* __restrict__ Fb,
                                                     int64_t ldfb, int64_t nb, int nfeat, int64_t tilesA,
                                                     double* __restrict__ out, int64_t ld) {
    extern __shared__ double sm[];
    double* Fr = sm;
    double* Fc = sm + nfeat * 64;
    const int64_t bid = xcd_remap_c(blockIdx.x, gridDim.x);
    const int64_t ti = bid % tilesA, tj = bid / tilesA;
    const int t = threadIdx.x, i = t & 63, cg = t >> 6;
    for (int e = t; e < nfeat * 64; e += 256) {
        const int f = e >> 6, p = e & 63;
        const int64_t ra = ti * 64 + p, rb = tj * 64 + p;
        Fr[e] = (ra < na) ? Fa[ra + (int64_t)f * ldfa] : 0.0;
        Fc[e] = (rb < nb) ? Fb[rb + (int64_t)f * ldfb] : 0.0;
    }
    __syncthreads();
    const int64_t gi = ti * 64 + i;
#pragma unroll 2
    for (int q = 0; q < 16; ++q) {
        const int c = cg + 4 * q;
        const int64_t gj = tj * 64 + c;
        const double v = (gi < na && gj < nb) ? prog_value(P, Fr + i, 64, Fc + c, 64) : 0.0;
        out[gi + gj * ld] = v;
    }
}

void launch_build_cross(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* Fa, int64_t lda_f,
                        int64_t na, int64_t Na, const double* Fb, int64_t ldb_f, int64_t nb, int64_t Nb,
                        double* out, int64_t ld) {
    const int64_t ta = Na / 64, tb = Nb / 64;
    const int nfeat = hp.nfeat > 0 ? hp.nfeat : 1;
    const size_t shm = (size_t)nfeat * 64 * 2 * sizeof(double);
    hipLaunchKernelGGL(k_build_cross, dim3((unsigned)(ta * tb)), dim3(256), shm, s, prog, Fa, lda_f, na, Fb, ldb_f,
                       nb, nfeat, ta, out, ld);
}

__global__ void k_build_diag(const DevProg* __restrict__ P, const double* __restrict__ Fa, int64_t lda,
                             const double* __restrict__ Fb, int64_t ldb, int64_t n, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = prog_value(P, Fa + i, (int)lda, Fb + i, (int)ldb);
}

void launch_build_diag(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* Fa, int64_t lda_f,
                       const double* Fb, int64_t ldb_f, int64_t n, double* out) {
    (void)hp;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_build_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, prog, Fa, lda_f, Fb, ldb_f,
                       n, out);
}

// ---------------------------------------------------------------------------
// Matrix-free y = (K + jitter I) v in FP64 (recomputes K from the features).
// Grid: (row blocks) x (column chunks); partials are summed in chunk order.
// ---------------------------------------------------------------------------
int kmatvec_chunks(int64_t N) {
    const int64_t T = N / 64;
    int64_t c = (2048 + T - 1) / (T > 0 ? T : 1);
    if (c < 1) c = 1;
    if (c > T) c = T;
    return (int)c;
}

__global__ __launch_bounds__(256) void k_kmatvec(const DevProg* __restrict__ P, const double* __restrict__ F,
                                                 int64_t ldf, int nfeat, int64_t n, int64_t T, int nchunk,
                                                 const double* __restrict__ v, double* __restrict__ part,
                                                 int64_t sF, int64_t sV, const int64_t* __restrict__ nvec) {
    {
        const int64_t b = blockIdx.z;  // batch
        P += (nvec ? b : 0);
        F += b * sF;
        v += b * sV;
        part += b * (int64_t)nchunk * T * 64;
        if (nvec) n = nvec[b];
    }
    extern __shared__ double sm[];
    double* Fr = sm;
    double* Fc = sm + nfeat * 64;
    __shared__ double vc[64];
    __shared__ double red[256];
    const int64_t ti = blockIdx.x;
    const int ch = blockIdx.y;
    const int t = threadIdx.x, i = t & 63, cg = t >> 6;
    for (int e = t; e < nfeat * 64; e += 256) {
        const int f = e >> 6, p = e & 63;
        Fr[e] = F[ti * 64 + p + (int64_t)f * ldf];
    }
    const int64_t gi = ti * 64 + i;
    double acc = 0.0;
    const int64_t per = (T + nchunk - 1) / nchunk;
    const int64_t tb = ch * per, te = (tb + per < T) ? tb + per : T;
    for (int64_t tj = tb; tj < te; ++tj) {
        __syncthreads();
        for (int e = t; e < nfeat * 64; e += 256) {
            const int f = e >> 6, p = e & 63;
            Fc[e] = F[tj * 64 + p + (int64_t)f * ldf];
        }
        if (t < 64) vc[t] = (tj * 64 + t < n) ? v[tj * 64 + t] : 0.0;
        __syncthreads();
        if (gi < n) {
#pragma unroll 2
            for (int q = 0; q < 16; ++q) {
                const int c = cg + 4 * q;
                acc += prog_value(P, Fr + i, 64, Fc + c, 64) * vc[c];
            }
        }
    }
    red[t] = acc;
    __syncthreads();
    if (t < 64) part[(int64_t)ch * (T * 64) + gi] = red[t] + red[t + 64] + red[t + 128] + red[t + 192];
}

__global__ void k_kmatvec_fin(const double* __restrict__ part, int64_t N, int nchunk, int64_t n, double jitter,
                              const double* __restrict__ v, double* __restrict__ y, int64_t sV,
                              const int64_t* __restrict__ nvec) {
    {
        const int64_t b = blockIdx.y;
        part += b * (int64_t)nchunk * N;
        v += b * sV;
        y += b * sV;
        if (nvec) n = nvec[b];
    }
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * N + i];
    y[i] = (i < n) ? s + jitter * v[i] : 0.0;
}

void launch_kmatvec(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F, int64_t n, int64_t N,
                    double jitter, const double* v, double* y, double* part, int nchunk) {
    const int64_t T = N / 64;
    const int nfeat = hp.nfeat > 0 ? hp.nfeat : 1;
    const size_t shm = (size_t)nfeat * 64 * 2 * sizeof(double);
    hipLaunchKernelGGL(k_kmatvec, dim3((unsigned)T, (unsigned)nchunk), dim3(256), shm, s, prog, F, N, nfeat, n, T,
                       nchunk, v, part, (int64_t)0, (int64_t)0, (const int64_t*)nullptr);
    hipLaunchKernelGGL(k_kmatvec_fin, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, part, N, nchunk, n, jitter,
                       v, y, (int64_t)0, (const int64_t*)nullptr);
}

// batch: part must hold batch * nchunk * N doubles
void launch_kmatvec_batched(hipStream_t s, const DevProg* progs, int nfeat, const double* F, int64_t sF,
                            const int64_t* nvec, int64_t N, double jitter, const double* v, double* y, int64_t sV,
                            double* part, int nchunk, int batch) {
    const int64_t T = N / 64;
    nfeat = nfeat > 0 ? nfeat : 1;
    const size_t shm = (size_t)nfeat * 64 * 2 * sizeof(double);
    hipLaunchKernelGGL(k_kmatvec, dim3((unsigned)T, (unsigned)nchunk, (unsigned)batch), dim3(256), shm, s, progs, F,
                       N, nfeat, (int64_t)0, T, nchunk, v, part, sF, sV, nvec);
    hipLaunchKernelGGL(k_kmatvec_fin, dim3((unsigned)((N + 255) / 256), (unsigned)batch), dim3(256), 0, s, part, N,
                       nchunk, (int64_t)0, jitter, v, y, sV, nvec);
}

// ---------------------------------------------------------------------------
// Fused gradient reduction over lower tiles:
//   part[tile][p] = sum_{(i,j) in tile, i>=j} w_ij (a_i a_j - c W_ij) dK_ij/dlog(theta_p)
// with w = 2 off the diagonal and 1 on it.  Fast path: one stationary term
// (SEXP/M32/M52) with up to NF features, all P accumulators in registers.
// ---------------------------------------------------------------------------
template <typename WT, int KIND, int NF>
__device__ __forceinline__ void grad_stationary_tile(const DevTerm& T, const double* Fr, const double* Fc,
                                                     const WT* __restrict__ Wt, int64_t ldw, double wscale,
                                                     const double* ar, const double* ac, int64_t gi0, int64_t gj0,
                                                     int64_t n, bool diag_tile, double (&acc)[NF + 1]) {
    const int t = threadIdx.x, i = t & 63, cg = t >> 6;
    const int64_t gi = gi0 + i;
    if (gi >= n) return;
    double xi[NF], il2[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        xi[f] = (f < T.nf) ? Fr[(T.f0 + f) * 64 + i] : 0.0;
        il2[f] = (f < T.nf) ? T.il2[f] : 0.0;
    }
    const double ai = ar[i];
#pragma unroll 2
    for (int q = 0; q < 16; ++q) {
        const int c = cg + 4 * q;
        const int64_t gj = gj0 + c;
        if (gj >= n || (diag_tile && c > i)) continue;
        const double w = (diag_tile && c == i) ? 1.0 : 2.0;
        const double coef = w * (ai * ac[c] - wscale * (double)Wt[i + (int64_t)c * ldw]);
        double t2[NF];
        double rho = 0.0;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const double dd = xi[f] - ((f < T.nf) ? Fc[(T.f0 + f) * 64 + c] : 0.0);
            t2[f] = dd * dd * il2[f];
            rho += t2[f];
        }
        double g, dg;
        stat_gd(KIND, rho, T.m2, g, dg);
        const double cd = -2.0 * coef * dg;
#pragma unroll
        for (int f = 0; f < NF; ++f) acc[f] += cd * t2[f];
        acc[NF] += 2.0 * coef * g;
    }
}

template <typename WT, int KIND, int NF>
__global__ __launch_bounds__(256) void k_grad_fast(const DevProg* __restrict__ P, const double* __restrict__ F,
                                                   int64_t ldf, int nfeat, int64_t n, const WT* __restrict__ W,
                                                   int64_t ldw, double wscale, const double* __restrict__ alpha,
                                                   double* __restrict__ part, int ngp, int64_t sF, int64_t sW,
                                                   int64_t sV, const int64_t* __restrict__ nvec,
                                                   const double* __restrict__ wsc) {
    {
        const int64_t b = blockIdx.y;  // batch
        P += (nvec ? b : 0);
        F += b * sF;
        W += b * sW;
        alpha += b * sV;
        part += b * (int64_t)gridDim.x * ngp;
        if (nvec) n = nvec[b];
        if (wsc) wscale = wsc[b];
    }
    extern __shared__ double sm[];
    double* Fr = sm;
    double* Fc = sm + nfeat * 64;
    __shared__ double ar[64], ac[64];
    __shared__ double red[256];
    const int64_t tile = blockIdx.x;
    int64_t ti, tj;
    lower_tile_c(tile, ti, tj);
    const int t = threadIdx.x;
    for (int e = t; e < nfeat * 64; e += 256) {
        const int f = e >> 6, p = e & 63;
        Fr[e] = F[ti * 64 + p + (int64_t)f * ldf];
        Fc[e] = F[tj * 64 + p + (int64_t)f * ldf];
    }
    if (t < 64) {
        ar[t] = (ti * 64 + t < n) ? alpha[ti * 64 + t] : 0.0;
        ac[t] = (tj * 64 + t < n) ? alpha[tj * 64 + t] : 0.0;
    }
    __syncthreads();
    const DevTerm& T = P->t[0];
    double acc[NF + 1];
#pragma unroll
    for (int p = 0; p <= NF; ++p) acc[p] = 0.0;
    grad_stationary_tile<WT, KIND, NF>(T, Fr, Fc, W + ti * 64 + tj * 64 * ldw, ldw, wscale, ar, ac, ti * 64, tj * 64,
                                       n, ti == tj, acc);
    // deterministic block reduction, one parameter at a time
    // slots: lscale features [0, nl) (iso: summed into slot 0), magn at nl
    const int nl = T.nl;
    for (int p = 0; p < ngp; ++p) {
        double v;
        if (p == nl) v = acc[NF];
        else if (nl == 1) {
            v = 0.0;
#pragma unroll
            for (int f = 0; f < NF; ++f) v += acc[f];
        } else {
            v = 0.0;
#pragma unroll
            for (int f = 0; f < NF; ++f) v = (f == p) ? acc[f] : v;
        }
        red[t] = v;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o) red[t] += red[t + o];
            __syncthreads();
        }
        if (t == 0) part[tile * ngp + p] = red[0];
        __syncthreads();
    }
}

// generic path: one parameter at a time, any program
template <typename WT>
__global__ __launch_bounds__(256) void k_grad_generic(const DevProg* __restrict__ P, const double* __restrict__ F,
                                                      int64_t ldf, int nfeat, int64_t n, const WT* __restrict__ W,
                                                      int64_t ldw, double wscale, const double* __restrict__ alpha,
                                                      double* __restrict__ part, int ngp, int64_t sF, int64_t sW,
                                                      int64_t sV, const int64_t* __restrict__ nvec,
                                                      const double* __restrict__ wsc) {
    {
        const int64_t b = blockIdx.y;  // batch
        P += (nvec ? b : 0);
        F += b * sF;
        W += b * sW;
        alpha += b * sV;
        part += b * (int64_t)gridDim.x * ngp;
        if (nvec) n = nvec[b];
        if (wsc) wscale = wsc[b];
    }
    extern __shared__ double sm[];
    double* Fr = sm;
    double* Fc = sm + nfeat * 64;
    __shared__ double ar[64], ac[64];
    __shared__ double red[256];
    const int64_t tile = blockIdx.x;
    int64_t ti, tj;
    lower_tile_c(tile, ti, tj);
    const int t = threadIdx.x, i = t & 63, cg = t >> 6;
    for (int e = t; e < nfeat * 64; e += 256) {
        const int f = e >> 6, p = e & 63;
        Fr[e] = F[ti * 64 + p + (int64_t)f * ldf];
        Fc[e] = F[tj * 64 + p + (int64_t)f * ldf];
    }
    if (t < 64) {
        ar[t] = (ti * 64 + t < n) ? alpha[ti * 64 + t] : 0.0;
        ac[t] = (tj * 64 + t < n) ? alpha[tj * 64 + t] : 0.0;
    }
    __syncthreads();
    const WT* Wt = W + ti * 64 + tj * 64 * ldw;
    const int64_t gi = ti * 64 + i;
    double coef[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = cg + 4 * q;
        const int64_t gj = tj * 64 + c;
        const bool ok = gi < n && gj < n && !(ti == tj && c > i);
        const double w = (ti == tj && c == i) ? 1.0 : 2.0;
        coef[q] = ok ? w * (ar[i] * ac[c] - wscale * (double)Wt[i + (int64_t)c * ldw]) : 0.0;
    }
    for (int p = 0; p < ngp; ++p) {
        double v = 0.0;
#pragma unroll 2
        for (int q = 0; q < 16; ++q) {
            if (coef[q] != 0.0) v += coef[q] * prog_dparam(P, p, Fr + i, 64, Fc + cg + 4 * q, 64);
        }
        red[t] = v;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o) red[t] += red[t + o];
            __syncthreads();
        }
        if (t == 0) part[tile * ngp + p] = red[0];
        __syncthreads();
    }
}

template <typename WT>
static void grad_launch(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F, int64_t n,
                        int64_t N, const WT* W, int64_t ldw, double wscale, const double* alpha, double* part,
                        int batch = 1, int64_t sF = 0, int64_t sW = 0, int64_t sV = 0, const int64_t* nvec = nullptr,
                        const double* wsc = nullptr) {
    const int64_t T = N / 64;
    const int nfeat = hp.nfeat > 0 ? hp.nfeat : 1;
    const size_t shm = (size_t)nfeat * 64 * 2 * sizeof(double);
    dim3 grid((unsigned)(T * (T + 1) / 2), (unsigned)batch), block(256);
    const int ngp = hp.ngrad;
    const DevTerm& t0 = hp.t[0];
    const bool fast = hp.nterms == 1 &&
                      (t0.kind == GPL_CF_SEXP || t0.kind == GPL_CF_MATERN32 || t0.kind == GPL_CF_MATERN52) &&
                      t0.nf <= 16 && t0.f0 == 0;
    if (fast) {
        const int nfp = t0.nf <= 1 ? 1 : t0.nf <= 2 ? 2 : t0.nf <= 4 ? 4 : t0.nf <= 8 ? 8 : 16;
#define GPL_GRAD_FAST(KD, NFv)                                                                                   \
    hipLaunchKernelGGL((k_grad_fast<WT, KD, NFv>), grid, block, shm, s, prog, F, N, nfeat, n, W, ldw, wscale, alpha, \
                       part, ngp, sF, sW, sV, nvec, wsc)
#define GPL_GRAD_NF(KD)                            \
    switch (nfp) {                                 \
        case 1: GPL_GRAD_FAST(KD, 1); break;       \
        case 2: GPL_GRAD_FAST(KD, 2); break;       \
        case 4: GPL_GRAD_FAST(KD, 4); break;       \
        case 8: GPL_GRAD_FAST(KD, 8); break;       \
        default: GPL_GRAD_FAST(KD, 16); break;     \
    }
        if (t0.kind == GPL_CF_SEXP) { GPL_GRAD_NF(GPL_CF_SEXP) }
        else if (t0.kind == GPL_CF_MATERN32) { GPL_GRAD_NF(GPL_CF_MATERN32) }
        else { GPL_GRAD_NF(GPL_CF_MATERN52) }
#undef GPL_GRAD_NF
#undef GPL_GRAD_FAST
    } else {
        hipLaunchKernelGGL(k_grad_generic<WT>, grid, block, shm, s, prog, F, N, nfeat, n, W, ldw, wscale, alpha,
                           part, ngp, sF, sW, sV, nvec, wsc);
    }
}

void launch_grad_tiles(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F, int64_t n, int64_t N,
                       const double* W, int64_t ldw, double wscale, const double* alpha, double* part) {
    grad_launch<double>(s, prog, hp, F, n, N, W, ldw, wscale, alpha, part);
}

void launch_grad_tiles_f32(hipStream_t s, const DevProg* prog, const DevProg& hp, const double* F, int64_t n,
                           int64_t N, const float* W, int64_t ldw, double wscale, const double* alpha,
                           double* part) {
    grad_launch<float>(s, prog, hp, F, n, N, W, ldw, wscale, alpha, part);
}

// batch (all problems share the term structure of hp): part holds batch * ntiles * ngrad
void launch_grad_tiles_f32_batched(hipStream_t s, const DevProg* progs, const DevProg& hp, const double* F,
                                   int64_t sF, const int64_t* nvec, int64_t N, const float* W, int64_t sW,
                                   const double* wsc, const double* alpha, int64_t sV, double* part, int batch) {
    grad_launch<float>(s, progs, hp, F, 0, N, W, N, 1.0, alpha, part, batch, sF, sW, sV, nvec, wsc);
}

}  // namespace gpl
