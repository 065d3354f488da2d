// Elementwise likelihood work and deterministic reductions for gfx950.
// Restates R/lik.R: pseudo-data for Laplace (:188-201) with the reference's
// forward-difference curvature (:449-454), log densities (:297-366), first
// derivatives (:376-443), responses (:460-498), and the EP site update with
// Gauss-Hermite tilted moments (:207-270, R/quadrature.R, logsumexp
// R/gp_loo.R:136-144).
#include "gpl_internal.h"

#include <cmath>

namespace gpl {

constexpr double kLog2Pi = 1.8378770664093454836;
constexpr double kInvSqrt2Pi = 0.39894228040143267794;

__device__ __forceinline__ double d_response(int lik, int link, double f) {
    if (lik == GPL_LIK_POISSON) return exp(f);
    if (link == GPL_LINK_PROBIT) return 0.5 * erfc(-f * 0.70710678118654752440);
    return 1.0 / (1.0 + exp(-f));
}

// digamma(x) for x > 0: the recurrence psi(x) = psi(x + 1) - 1/x raises the
// argument to at least 10, then the asymptotic series
// psi(x) ~ ln x - 1/(2x) - sum_k B_2k / (2k x^2k) (Abramowitz & Stegun 6.3.18)
// through the x^-12 term; the first omitted term is below 1e-16 at x = 10.
__device__ double d_digamma(double x) {
    double acc = 0.0;
    while (x < 10.0) {
        acc -= 1.0 / x;
        x += 1.0;
    }
    const double r = 1.0 / x, r2 = r * r;
    // B_2k / (2k) for k = 1..6: 1/12, -1/120, 1/252, -1/240, 1/132, -691/32760
    const double series =
        r2 * (1.0 / 12 - r2 * (1.0 / 120 - r2 * (1.0 / 252 - r2 * (1.0 / 240 - r2 * (1.0 / 132 - r2 * (691.0 / 32760))))));
    return acc + log(x) - 0.5 * r - series;
}

// d loglik / d f at f (offset already added); par is the beta-binomial phi
__device__ __forceinline__ double d_loglik_d(int lik, int link, double par, double f, double y, double tr) {
    const double mu = d_response(lik, link, f);
    if (lik == GPL_LIK_POISSON) return (y / mu - 1.0) * mu;
    const double dmu = (link == GPL_LINK_PROBIT) ? kInvSqrt2Pi * exp(-0.5 * f * f) : mu * (1.0 - mu);
    if (lik == GPL_LIK_BETABINOM) {  // R/lik.R:400-425
        const double a = mu / par, b = (1.0 - mu) / par;
        return (dmu / par) * (d_digamma(y + a) - d_digamma(tr - y + b)) + (dmu / par) * (-d_digamma(a) + d_digamma(b));
    }
    return dmu * (y / mu - (tr - y) / (1.0 - mu));
}

// log p(y | f); par is the Gaussian sigma or the beta-binomial phi
__device__ __forceinline__ double d_loglik(int lik, int link, double par, double f, double y, double tr) {
    if (lik == GPL_LIK_GAUSSIAN) {
        const double u = (y - f) / par;
        return -0.5 * kLog2Pi - log(par) - 0.5 * u * u;
    }
    const double mu = d_response(lik, link, f);
    if (lik == GPL_LIK_POISSON) return (y > 0.0 ? y * log(mu) : 0.0) - mu - lgamma(y + 1.0);
    const double lc = lgamma(tr + 1.0) - lgamma(y + 1.0) - lgamma(tr - y + 1.0);
    if (lik == GPL_LIK_BETABINOM) {  // R/lik.R:332-352
        const double a = mu / par, b = (1.0 - mu) / par;
        const double t2 = lgamma(y + a) + lgamma(tr - y + b) - lgamma(tr + a + b);
        const double t3 = lgamma(a + b) - lgamma(a) - lgamma(b);
        return lc + t2 + t3;
    }
    return lc + (y > 0.0 ? y * log(mu) : 0.0) + (tr - y > 0.0 ? (tr - y) * log1p(-mu) : 0.0);
}

// Exact second and third derivatives of log p(y | f) in f (offset added) for the
// log-concave likelihoods (not in the reference, whose curvature is the forward
// difference of R/lik.R:449-454).  logit: d2 = -t mu (1-mu), d3 = d2 (1 - 2 mu);
// probit, with r1 = phi/Phi and r0 = phi/(1-Phi) (both tails from erfc):
// d2 = -y r1 (f + r1) - (t-y) r0 (r0 - f),
// d3 = -y r1 [1 - (f + r1)(f + 2 r1)] - (t-y) r0 [(r0 - f)(2 r0 - f) - 1];
// Poisson (log link): d2 = d3 = -exp(f); Gaussian: -1/sigma^2 and 0.
__device__ __forceinline__ void d_loglik_d23(int lik, int link, double par, double f, double y, double tr, double& d2,
                                             double& d3) {
    if (lik == GPL_LIK_GAUSSIAN) {
        d2 = -1.0 / (par * par);
        d3 = 0.0;
        return;
    }
    if (lik == GPL_LIK_POISSON) {
        d2 = d3 = -exp(f);
        return;
    }
    if (link == GPL_LINK_PROBIT) {
        const double phi = kInvSqrt2Pi * exp(-0.5 * f * f);
        d2 = d3 = 0.0;
        if (y > 0.0) {
            const double r1 = phi / (0.5 * erfc(-f * 0.70710678118654752440));
            d2 -= y * r1 * (f + r1);
            d3 -= y * r1 * (1.0 - (f + r1) * (f + 2.0 * r1));
        }
        if (tr - y > 0.0) {
            const double r0 = phi / (0.5 * erfc(f * 0.70710678118654752440));
            d2 -= (tr - y) * r0 * (r0 - f);
            d3 -= (tr - y) * r0 * ((r0 - f) * (2.0 * r0 - f) - 1.0);
        }
        return;
    }
    const double mu = 1.0 / (1.0 + exp(-f)), nu = 1.0 / (1.0 + exp(f));  // nu = 1 - mu without cancellation
    d2 = -tr * mu * nu;
    d3 = d2 * (nu - mu);
}

__global__ void k_pseudo_la(int lik, int link, double sigma, int64_t n, const double* __restrict__ f,
                            const double* __restrict__ y, const double* __restrict__ trials,
                            const double* __restrict__ offset, double* __restrict__ z, double* __restrict__ V,
                            double* __restrict__ sv, int* __restrict__ bad, int exact) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (lik == GPL_LIK_GAUSSIAN) {  // R/lik.R:188-192: z = y (offset ignored), V = sigma^2
        z[i] = y[i];
        V[i] = sigma * sigma;
        sv[i] = sqrt(1.0 / (sigma * sigma));
        return;
    }
    const double off = offset ? offset[i] : 0.0;
    const double tr = trials ? trials[i] : 1.0;
    const double eps = 1e-6;
    const double fi = f[i];
    const double g = d_loglik_d(lik, link, sigma, fi + off, y[i], tr);
    double g2;
    if (exact) {
        double d3;
        d_loglik_d23(lik, link, sigma, fi + off, y[i], tr, g2, d3);
    } else {
        const double ge = d_loglik_d(lik, link, sigma, (fi + eps) + off, y[i], tr);
        g2 = (ge - g) / eps;
    }
    const double sg = (g2 > 0.0) ? 1.0 : ((g2 < 0.0) ? -1.0 : 0.0);
    g2 = sg * fmax(fabs(g2), 1e-12);
    const double Vi = -1.0 / g2;
    z[i] = fi - g / g2;
    V[i] = Vi;
    sv[i] = sqrt(1.0 / Vi);
    if (isnan(Vi) || isnan(z[i])) atomicOr(bad, 2);
    else if (!(Vi > 0.0)) atomicOr(bad, 4);
}

void lik_pseudodata_la(hipStream_t s, int lik, int link, double sigma, int64_t n, const double* f, const double* y,
                       const double* trials, const double* offset, double* z, double* V, double* sv, int* bad,
                       int exact) {
    hipLaunchKernelGGL(k_pseudo_la, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lik, link, sigma, n, f, y,
                       trials, offset, z, V, sv, bad, exact);
}

// Per-point vectors of the Laplace gradient at the mode f (R&W Algorithm 5.1):
// u = d log p / d f, and s2 = d logZ / d f = 1/2 [(K^-1 + W)^-1]_ii d3_i with
// (K^-1 + W)^-1 = W^-1/2 (I - B^-1) W^-1/2, so its diagonal is V_i (1 - B^-1_ii)
// (B^-1 lower tiles, leading dimension ldb).  Padding (i >= n) gets zeros.
__global__ void k_la_grad_vec(int lik, int link, double par, int64_t n, int64_t N, const double* __restrict__ f,
                              const double* __restrict__ y, const double* __restrict__ trials,
                              const double* __restrict__ offset, const double* __restrict__ V,
                              const double* __restrict__ Binv, int64_t ldb, double* __restrict__ u,
                              double* __restrict__ s2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (i >= n) {
        u[i] = 0.0;
        s2[i] = 0.0;
        return;
    }
    const double fi = f[i] + (offset ? offset[i] : 0.0);
    const double tr = trials ? trials[i] : 1.0;
    double d2, d3;
    d_loglik_d23(lik, link, par, fi, y[i], tr, d2, d3);
    u[i] = d_loglik_d(lik, link, par, fi, y[i], tr);
    s2[i] = 0.5 * V[i] * (1.0 - Binv[i + i * ldb]) * d3;
}

void lik_la_grad_vectors(hipStream_t s, int lik, int link, double par, int64_t n, int64_t N, const double* f,
                         const double* y, const double* trials, const double* offset, const double* V,
                         const double* Binv, int64_t ldb, double* u, double* s2) {
    hipLaunchKernelGGL(k_la_grad_vec, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, lik, link, par, n, N, f, y,
                       trials, offset, V, Binv, ldb, u, s2);
}

__global__ void k_loglik(int lik, int link, double sigma, int64_t n, const double* __restrict__ f,
                         const double* __restrict__ y, const double* __restrict__ trials,
                         const double* __restrict__ offset, double* __restrict__ ll) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double off = offset ? offset[i] : 0.0;
    const double tr = trials ? trials[i] : 1.0;
    ll[i] = d_loglik(lik, link, sigma, f[i] + off, y[i], tr);
}

void lik_loglik(hipStream_t s, int lik, int link, double sigma, int64_t n, const double* f, const double* y,
                const double* trials, const double* offset, double* ll) {
    hipLaunchKernelGGL(k_loglik, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lik, link, sigma, n, f, y,
                       trials, offset, ll);
}

// EP site update (one thread per site, q <= 32 quadrature nodes)
__global__ void k_ep_sites(int lik, int link, double par, int64_t n, const double* __restrict__ fmean,
                           const double* __restrict__ fvar, const double* __restrict__ z_old,
                           const double* __restrict__ P_old, const double* __restrict__ y,
                           const double* __restrict__ trials, const double* __restrict__ offset, int q,
                           const double* __restrict__ gx, const double* __restrict__ gw, double damping,
                           double* __restrict__ z, double* __restrict__ V, double* __restrict__ sv,
                           double* __restrict__ logZs, double* __restrict__ cav_m, double* __restrict__ cav_v,
                           int* __restrict__ nan_flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double post_prec = 1.0 / fvar[i];
    const double Po = P_old[i], zo = z_old[i];
    const double cprec = post_prec - Po;
    const double cmean = 1.0 / cprec * (fmean[i] * post_prec - zo * Po);
    const double scale = sqrt(1.0 / cprec);
    const double off = offset ? offset[i] : 0.0;
    const double tr = trials ? trials[i] : 1.0;
    const double yi = y[i];
    double ll[32], fg[32];
    bool ok = true;
    double m = -INFINITY;
    for (int k = 0; k < q; ++k) {
        fg[k] = cmean + scale * gx[k];
        double v = d_loglik(lik, link, par, fg[k] + off, yi, tr);
        if (isnan(v)) { v = -INFINITY; ok = false; }
        ll[k] = v;
        m = fmax(m, v);
    }
    double sw = 0.0;
    for (int k = 0; k < q; ++k) sw += gw[k] * exp(ll[k] - m);
    const double logC = m + log(sw);
    const double C = exp(logC);
    double mt = 0.0;
    for (int k = 0; k < q; ++k) mt += fg[k] * exp(ll[k]) * gw[k];
    mt = (1.0 / C) * mt;
    double vt = 0.0;
    for (int k = 0; k < q; ++k) {
        const double d = fg[k] - mt;
        vt += d * d * exp(ll[k]) * gw[k];
    }
    vt = (1.0 / C) * vt;
    const double cvar = 1.0 / cprec;
    double Pn = 1.0 / vt - 1.0 / cvar;
    double nun = mt / vt - cmean / cvar;
    const double nuo = zo * Po;
    nun = damping * nun + (1.0 - damping) * nuo;
    Pn = damping * Pn + (1.0 - damping) * Po;
    const double Vi = 1.0 / Pn;
    const double zi = (1.0 / Pn) * nun;
    z[i] = zi;
    V[i] = Vi;
    sv[i] = sqrt(1.0 / Vi);
    logZs[i] = logC;
    cav_m[i] = cmean;
    cav_v[i] = cvar;
    if (!ok) atomicOr(nan_flag, 1);
    if (isnan(Vi) || isnan(zi)) atomicOr(nan_flag, 2);
    else if (!(Vi > 0.0)) atomicOr(nan_flag, 4);
}

void ep_sites(hipStream_t s, int lik, int link, double par, int64_t n, const double* fmean, const double* fvar,
              const double* z_old, const double* P_old, const double* y, const double* trials, const double* offset,
              int q, const double* gh_x, const double* gh_w, double damping, double* z, double* V, double* sv,
              double* logZs, double* cav_m, double* cav_v, int* nan_flag) {
    hipLaunchKernelGGL(k_ep_sites, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, lik, link, par, n, fmean, fvar,
                       z_old, P_old, y, trials, offset, q, gh_x, gh_w, damping, z, V, sv, logZs, cav_m, cav_v,
                       nan_flag);
}

// fvar_i = V_i (1 - [B^-1]_ii); per-site evidence term
__global__ void k_ep_post(int64_t n, const double* __restrict__ V, const double* __restrict__ z,
                          const double* __restrict__ binv_diag, const double* __restrict__ cav_m,
                          const double* __restrict__ cav_v, const double* __restrict__ logZs,
                          double* __restrict__ fvar, double* __restrict__ term) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double Vi = V[i];
    fvar[i] = Vi * (1.0 - binv_diag[i]);
    const double cv = cav_v[i] + Vi;
    const double dz = cav_m[i] - z[i];
    term[i] = logZs[i] + 0.5 * log(cv) + 0.5 * dz * dz / cv - 0.5 * log(Vi);
}

void ep_post(hipStream_t s, int64_t n, const double* V, const double* z, const double* binv_diag, const double* cav_m,
             const double* cav_v, const double* logZs, double* fvar, double* term) {
    hipLaunchKernelGGL(k_ep_post, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, V, z, binv_diag, cav_m,
                       cav_v, logZs, fvar, term);
}

__global__ void k_axpby(int64_t n, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                        double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * x[i] + (y ? b * y[i] : 0.0);
}

void vec_axpby(hipStream_t s, int64_t n, double a, const double* x, double b, const double* y, double* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, x, b, y, out);
}

__global__ void k_mul(int64_t n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] * b[i];
}

void vec_mul(hipStream_t s, int64_t n, const double* a, const double* b, double* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mul, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, b, out);
}

__global__ void k_recip(int64_t n, const double* __restrict__ a, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = 1.0 / a[i];
}

void vec_recip(hipStream_t s, int64_t n, const double* a, double* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_recip, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, out);
}

// ---------------------------------------------------------------------------
// Deterministic reductions: one workgroup, fixed per-thread order, fixed tree.
// ---------------------------------------------------------------------------
template <int OP>  // 0 sum, 1 dot, 2 max |a-b|
__global__ __launch_bounds__(1024) void k_reduce(const double* __restrict__ a, const double* __restrict__ b,
                                                 int64_t n, double* __restrict__ out) {
    __shared__ double red[1024];
    const int t = threadIdx.x;
    double s = (OP == 2) ? 0.0 : 0.0;
    for (int64_t i = t; i < n; i += 1024) {
        if (OP == 0) s += a[i];
        else if (OP == 1) s += a[i] * b[i];
        else {
            const double d = fabs(a[i] - b[i]);
            s = (d > s || isnan(d)) ? d : s;
        }
    }
    red[t] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o) {
            if (OP == 2) {
                const double u = red[t + o];
                red[t] = (u > red[t] || isnan(u)) ? u : red[t];
            } else {
                red[t] += red[t + o];
            }
        }
        __syncthreads();
    }
    if (t == 0) out[0] = red[0];
}

void reduce_sum(hipStream_t s, const double* v, int64_t n, double* out) {
    hipLaunchKernelGGL(k_reduce<0>, dim3(1), dim3(1024), 0, s, v, v, n, out);
}
void reduce_dot(hipStream_t s, const double* a, const double* b, int64_t n, double* out) {
    hipLaunchKernelGGL(k_reduce<1>, dim3(1), dim3(1024), 0, s, a, b, n, out);
}
void reduce_maxabsdiff(hipStream_t s, const double* a, const double* b, int64_t n, double* out) {
    hipLaunchKernelGGL(k_reduce<2>, dim3(1), dim3(1024), 0, s, a, b, n, out);
}

__global__ __launch_bounds__(256) void k_reduce_cols(const double* __restrict__ part, int64_t rows, int64_t ld,
                                                     double* __restrict__ out) {
    __shared__ double red[256];
    const int64_t c = blockIdx.x;
    const int t = threadIdx.x;
    double s = 0.0;
    for (int64_t r = t; r < rows; r += 256) s += part[r * ld + c];
    red[t] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) out[c] = red[0];
}

void reduce_cols(hipStream_t s, const double* part, int64_t rows, int64_t ld, int64_t ncols, double* out) {
    if (ncols <= 0) return;
    hipLaunchKernelGGL(k_reduce_cols, dim3((unsigned)ncols), dim3(256), 0, s, part, rows, ld, out);
}

}  // namespace gpl

namespace gpl {

// ---------------------------------------------------------------------------
// Segmented (one segment per batched problem, stride N) helpers
// ---------------------------------------------------------------------------
// out[b*N + i] = (i < n_b) ? scal[b] * (in ? in[b*N + i] : 1) : 0
__global__ void k_seg_scale(const double* __restrict__ in, const double* __restrict__ scal,
                            const int64_t* __restrict__ nvec, int64_t N, double* __restrict__ out) {
    const int64_t b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t o = b * N + i;
    out[o] = (i < nvec[b]) ? scal[b] * (in ? in[o] : 1.0) : 0.0;
}

void seg_scale(hipStream_t s, const double* in, const double* scal, const int64_t* nvec, int64_t N, double* out,
               int batch) {
    hipLaunchKernelGGL(k_seg_scale, dim3((unsigned)((N + 255) / 256), (unsigned)batch), dim3(256), 0, s, in, scal,
                       nvec, N, out);
}

// per segment: out[b*ostride + slot] = sum_i a_i (b == null) or sum_i a_i b_i over i < n_b
__global__ __launch_bounds__(256) void k_seg_dot(const double* __restrict__ a, const double* __restrict__ bv,
                                                 const int64_t* __restrict__ nvec, int64_t N,
                                                 double* __restrict__ out, int ostride, int slot) {
    __shared__ double red[256];
    const int64_t b = blockIdx.x;
    const int t = threadIdx.x;
    const int64_t n = nvec[b];
    const double* pa = a + b * N;
    const double* pb = bv ? bv + b * N : nullptr;
    double s = 0.0;
    for (int64_t i = t; i < n; i += 256) s += pb ? pa[i] * pb[i] : pa[i];
    red[t] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) out[b * ostride + slot] = red[0];
}

void seg_dot(hipStream_t s, const double* a, const double* b, const int64_t* nvec, int64_t N, double* out,
             int ostride, int slot, int batch) {
    hipLaunchKernelGGL(k_seg_dot, dim3((unsigned)batch), dim3(256), 0, s, a, b, nvec, N, out, ostride, slot);
}

// out[b*ostride + slot0 + p] = sum_t part[(b*ntiles + t)*ng + p]
__global__ __launch_bounds__(256) void k_seg_cols(const double* __restrict__ part, int64_t ntiles, int ng,
                                                  double* __restrict__ out, int ostride, int slot0) {
    __shared__ double red[256];
    const int64_t b = blockIdx.y;
    const int p = blockIdx.x;
    const int t = threadIdx.x;
    const double* pp = part + b * ntiles * ng;
    double s = 0.0;
    for (int64_t r = t; r < ntiles; r += 256) s += pp[r * ng + p];
    red[t] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) out[b * ostride + slot0 + p] = red[0];
}

void seg_cols(hipStream_t s, const double* part, int64_t ntiles, int ng, double* out, int ostride, int slot0,
              int batch) {
    hipLaunchKernelGGL(k_seg_cols, dim3((unsigned)ng, (unsigned)batch), dim3(256), 0, s, part, ntiles, ng, out, ostride,
                       slot0);
}

// per problem b: out[b] = max_i |d_i| / max_i |u_i| over its i < n_b (the last
// refinement correction relative to the solution); NaN when any entry is not a number
__global__ __launch_bounds__(256) void k_seg_relmax(const double* __restrict__ d, const double* __restrict__ u,
                                                    const int64_t* __restrict__ nvec, int64_t N,
                                                    double* __restrict__ out) {
    __shared__ double rd[256], ru[256];
    __shared__ int bad[256];
    const int64_t b = blockIdx.x;
    const int t = threadIdx.x;
    const int64_t n = nvec[b];
    const double* pd = d + b * N;
    const double* pu = u + b * N;
    double a = 0.0, c = 0.0;
    int nan = 0;
    for (int64_t i = t; i < n; i += 256) {
        const double x = pd[i], y = pu[i];
        nan |= (x != x) | (y != y);
        a = fmax(a, fabs(x));
        c = fmax(c, fabs(y));
    }
    rd[t] = a;
    ru[t] = c;
    bad[t] = nan;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            rd[t] = fmax(rd[t], rd[t + o]);
            ru[t] = fmax(ru[t], ru[t + o]);
            bad[t] |= bad[t + o];
        }
        __syncthreads();
    }
    if (t == 0) out[b] = bad[0] ? __longlong_as_double(0x7ff8000000000000LL) : ((ru[0] > 0.0) ? rd[0] / ru[0] : rd[0]);
}

void seg_relmax(hipStream_t s, const double* d, const double* u, const int64_t* nvec, int64_t N, double* out,
                int batch) {
    hipLaunchKernelGGL(k_seg_relmax, dim3((unsigned)batch), dim3(256), 0, s, d, u, nvec, N, out);
}

// Per problem: log evidence and gradient from the segment sums.
// red per problem (stride rs): [0] y^T alpha, [1] sum log L_ii, [2] alpha^T alpha,
// [3] tr(B^-1), [4] first-order log|B| correction of the FP32 factor, [5..] gradient partials
__global__ void k_batch_finish(const double* __restrict__ red, int rs, const int64_t* __restrict__ nvec,
                               const double* __restrict__ sigma, const int32_t* __restrict__ idx, int ng, int nb,
                               double* __restrict__ logz, double* __restrict__ grad, int gstride) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const double* r = red + (int64_t)b * rs;
    const double n = (double)nvec[b];
    const double sg = sigma[b], s2 = sg * sg;
    const int p = idx[b];
    // log N(y | 0, K + sigma^2 I) with K + sigma^2 I = sigma^2 B:  -1/2 y^T alpha - 1/2 log|B| - n log sigma - n/2 log 2 pi
    logz[p] = -0.5 * r[0] - (r[1] + 0.5 * r[4]) - n * log(sg) - 0.5 * n * kLog2Pi;
    for (int q = 0; q < ng; ++q) grad[(int64_t)p * gstride + q] = -0.5 * r[5 + q];
    grad[(int64_t)p * gstride + ng] = -s2 * r[2] + r[3];
}

void batch_finish(hipStream_t s, const double* red, int rs, const int64_t* nvec, const double* sigma,
                  const int32_t* idx, int ng, int nb, double* logz, double* grad, int gstride) {
    hipLaunchKernelGGL(k_batch_finish, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, s, red, rs, nvec, sigma, idx, ng,
                       nb, logz, grad, gstride);
}

}  // namespace gpl
