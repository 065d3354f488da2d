/*
 * gplite_amd -- MI355X (gfx950) dense Gaussian-process core behind a C ABI.
 *
 * Every entry point takes plain host pointers and sizes.  Matrices are
 * column-major FP64 exactly as R hands them to `.Call` (x[i + j*n]).
 * Inputs are borrowed for the duration of the call; outputs are written
 * into caller-owned buffers.  Device-resident state (factors, alpha, the
 * inverse) lives in a gpl_model handle until gpl_model_free().
 *
 * All functions return a gpl_status; on failure gpl_last_error() gives a
 * message and gpl_last_error_minor() the failing leading minor (1-based) for
 * GPL_ERR_NOT_PD, mirroring R's "leading minor of order k is not positive".
 */
#ifndef GPLITE_AMD_H
#define GPLITE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GPL_OK = 0,
    GPL_ERR_ARG = 1,        /* invalid argument (R: stop()) */
    GPL_ERR_NOT_PD = 2,     /* Cholesky found a non-positive pivot (R: chol() error) */
    GPL_ERR_OOM = 3,        /* device allocation failed / admission refused */
    GPL_ERR_NOCONV = 4,     /* Newton / EP did not converge (R: warning or -Inf) */
    GPL_ERR_HIP = 5,        /* HIP runtime error */
    GPL_ERR_NAN = 6,        /* NaN in pseudo-data / quadrature (R: log_evidence = -Inf) */
    GPL_ERR_UNSUPPORTED = 7 /* a case the device path does not cover (named in the message) */
} gpl_status;

/* ---- covariance program ------------------------------------------------ */
/* Kinds follow the reference constructors cf_const ... cf_periodic
 * (R/cf.R:92-181).  A model's covariance is a list of terms: a term with
 * op == GPL_OP_SUM starts a new summand (R/cf.R:402-411, eval_cf.list), a term
 * with op == GPL_OP_PROD multiplies into the current summand (R/cf.R:491-497,
 * eval_cf.cf_prod).  The host applies `vars` selection and `normalize`
 * (R/misc.R:17-31) before the call; each term reads columns
 * [col0, col0+ncols) of the input matrix. */
enum {
    GPL_CF_CONST = 0, GPL_CF_LIN = 1, GPL_CF_SEXP = 2, GPL_CF_MATERN32 = 3,
    GPL_CF_MATERN52 = 4, GPL_CF_NN = 5, GPL_CF_PERIODIC = 6
};
enum { GPL_OP_SUM = 0, GPL_OP_PROD = 1 };

#define GPL_MAX_TERMS 8
#define GPL_MAX_COLS 16

typedef struct gpl_cf_term {
    int32_t kind;       /* GPL_CF_* */
    int32_t op;         /* GPL_OP_SUM / GPL_OP_PROD (ignored for the first term) */
    int32_t col0;       /* first input column read by this term */
    int32_t ncols;      /* number of input columns read (<= GPL_MAX_COLS) */
    int32_t base_kind;  /* GPL_CF_PERIODIC only: SEXP, MATERN32 or MATERN52 */
    int32_t n_lscale;   /* 1 (isotropic) or ncols (ARD); periodic: 1 or 2*ncols */
    double magn;
    double sigma0;      /* GPL_CF_NN */
    double sigma;       /* GPL_CF_NN */
    double period;      /* GPL_CF_PERIODIC */
    double lscale[2 * GPL_MAX_COLS];
} gpl_cf_term;

/* Gradient layout of one term (d energy / d log theta), in this order:
 *   CONST, LIN: magn
 *   SEXP, MATERN32, MATERN52: lscale[0..n_lscale-1], magn
 *   NN: sigma0, sigma, magn
 *   PERIODIC: period, lscale[0..n_lscale-1], magn
 * followed by the likelihood's parameters (Gaussian: sigma). */
int32_t gpl_grad_len(const gpl_cf_term* terms, int32_t nterms, int32_t lik_kind);

/* ---- likelihood and options -------------------------------------------- */
enum { GPL_LIK_GAUSSIAN = 0, GPL_LIK_BERNOULLI = 1, GPL_LIK_BINOMIAL = 2, GPL_LIK_POISSON = 3, GPL_LIK_BETABINOM = 4 };
enum { GPL_LINK_IDENTITY = 0, GPL_LINK_LOGIT = 1, GPL_LINK_PROBIT = 2, GPL_LINK_LOG = 3 };

typedef struct gpl_lik {
    int32_t kind;   /* GPL_LIK_* (R/lik.R:46-90) */
    int32_t link;   /* GPL_LINK_* */
    double sigma;   /* Gaussian noise sd */
    double phi;     /* beta-binomial overdispersion (lik_betabinom(phi), R/lik.R:74-80) */
} gpl_lik;

enum { GPL_APPROX_LAPLACE = 0, GPL_APPROX_EP = 1 };
enum { GPL_PREC_FP64 = 0, GPL_PREC_MIXED = 1 };

typedef struct gpl_opts {
    int32_t approx;       /* GPL_APPROX_LAPLACE (approx_laplace) / GPL_APPROX_EP (approx_ep) */
    int32_t maxiter;      /* approx_laplace(maxiter=30) / approx_ep(maxiter=100) */
    double tol;           /* approx_laplace(tol=1e-4); EP always uses 1e-4 (R/ep.R:286) */
    double damping;       /* approx_ep(damping=0.9) */
    double damping_min;   /* 0.1 (R/gp_fit.R:158) */
    int32_t quad_order;   /* approx_ep(quad_order=11) */
    int32_t precision;    /* GPL_PREC_FP64 (plain) or GPL_PREC_MIXED (FP32 MFMA + FP64 diagonal tiles and
                             refinement; Gaussian, Laplace and EP fits) */
    double jitter;        /* 1e-6 (R/gp.R:276-281) */
    int32_t refine_iters; /* mixed precision: max iterative-refinement sweeps (default 8) */
    int32_t flags;        /* GPL_FLAG_* bits */
} gpl_opts;

/* gpl_opts.flags: form L^-1 during a Gaussian fit (alpha then comes from two
 * triangular matrix-vector products, and a following gpl_energy_grad or
 * gpl_pred reuses it).  Set it when the gradient will be requested. */
#define GPL_FLAG_INVERSE 1
/* gpl_opts.flags: the Laplace curvature W = -d2 log p(y|f) in closed form
 * (logit, probit, log links; not the beta-binomial) in place of the reference's
 * forward difference with step 1e-6 (R/lik.R:449-454).  The evidence then
 * differs from the reference's by the O(1e-6) curvature error; gpl_energy_grad
 * needs it for a Laplace fit, because the exact gradient is that of the exact
 * Laplace approximation.  New (the reference has no such option). */
#define GPL_FLAG_ANALYTIC_W 2

void gpl_opts_default(gpl_opts* o);

/* ---- context ------------------------------------------------------------ */
typedef struct gpl_ctx gpl_ctx;
typedef struct gpl_model gpl_model;

int gpl_ctx_create(int32_t device, gpl_ctx** out);
void gpl_ctx_destroy(gpl_ctx* ctx);
const char* gpl_last_error(void);
int64_t gpl_last_error_minor(void);
int gpl_device_name(gpl_ctx* ctx, char* buf, int32_t len);

/* Per-kernel-class device timing with HIP events on the library's stream. */
int gpl_profile_enable(gpl_ctx* ctx, int32_t on);
int gpl_profile_reset(gpl_ctx* ctx);
/* Fills up to `cap` entries; returns the number of kernel classes. */
int gpl_profile_read(gpl_ctx* ctx, int32_t cap, char* names /* cap*32 bytes */,
                     double* total_ms, int64_t* launches, double* flops, double* bytes);

/* ---- Level 0: the four .Call routines of src/cf.cpp --------------------- */
/* Replace _gplite_cf_sexp_c / _matern32_c / _matern52_c (src/RcppExports.cpp:16-58,
 * src/cf.cpp:9-119) and _gplite_cf_nn_c (src/RcppExports.cpp:61-73,
 * src/cf.cpp:125-162).  x1 is n1 x d, x2 is n2 x d, column-major.  out is
 * n1 x n2 column-major, or min(n1,n2) x 1 when diag_only.  ARD: n_lscale == d. */
int gpl_cf_sexp(gpl_ctx* ctx, const double* x1, int64_t n1, const double* x2, int64_t n2,
                int64_t d, const double* lscale, int64_t n_lscale, double magn,
                int32_t diag_only, double* out);
int gpl_cf_matern32(gpl_ctx* ctx, const double* x1, int64_t n1, const double* x2, int64_t n2,
                    int64_t d, const double* lscale, int64_t n_lscale, double magn,
                    int32_t diag_only, double* out);
int gpl_cf_matern52(gpl_ctx* ctx, const double* x1, int64_t n1, const double* x2, int64_t n2,
                    int64_t d, const double* lscale, int64_t n_lscale, double magn,
                    int32_t diag_only, double* out);
int gpl_cf_nn(gpl_ctx* ctx, const double* x1, int64_t n1, const double* x2, int64_t n2,
              int64_t d, double sigma0, double sigma, double magn, int32_t diag_only,
              double* out);

/* eval_cf over a whole covariance program (R/cf.R:402-497). */
int gpl_eval_cf(gpl_ctx* ctx, const gpl_cf_term* terms, int32_t nterms,
                const double* x1, int64_t n1, const double* x2, int64_t n2, int64_t ncols,
                int32_t diag_only, double* out);

/* ---- Level 1: fused fit / energy / gradient / prediction --------------- */
/* gp_fit for method_full (R/gp_fit.R:54-60,91-99,150-161): builds
 * K = eval_cf(x,x) + jitter*I on the device, runs the Laplace Newton loop
 * (R/laplace.R:274-297; one step for a Gaussian likelihood) or parallel EP
 * (R/ep.R:286-330), and leaves the factor of B = I + S K S, alpha and the
 * posterior vectors resident in *inout.  If *inout is an existing model it is
 * refitted in place (its buffers are recycled when large enough); it must belong
 * to ctx (GPL_ERR_ARG otherwise).  NULL allocates a new model.  trials/offset
 * may be NULL. */
int gpl_fit(gpl_ctx* ctx, const gpl_cf_term* terms, int32_t nterms,
            const double* x, int64_t n, int64_t ncols, const double* y,
            const double* trials, const double* offset, const gpl_lik* lik,
            const gpl_opts* opts, gpl_model** inout, double* log_evidence);

/* d(-log evidence)/d(log theta) (new: the reference optimises without
 * gradients, R/gp_optim.R:94).  grad has gpl_grad_len() entries (full layout,
 * fixed parameters included).  Gaussian likelihood: any precision.  EP (FP64):
 * the fixed-point gradient -1/2 [a^T C_j a - tr((K + V)^-1 C_j)], a = (K + V)^-1 z
 * (Rasmussen & Williams 2006, eq. 5.27).  Laplace (FP64, GPL_FLAG_ANALYTIC_W):
 * R&W Algorithm 5.1, the explicit term plus the mode's implicit dependence
 * through d3 log p.  Latent likelihoods contribute no parameter of their own.
 * GPL_ERR_UNSUPPORTED for mixed-precision latent fits, split (non-log-concave)
 * fits and a Laplace fit without GPL_FLAG_ANALYTIC_W. */
int gpl_energy_grad(gpl_model* model, double* grad);

/* gp_pred posterior for method_full (R/gp_pred.R:202-232): latent mean
 * K_t alpha and (want_var) variance diag(K_tt) - colsum((C^-1 K_t^T)^2).
 * xt is m x ncols column-major. */
int gpl_pred(gpl_model* model, const double* xt, int64_t m, int32_t want_var,
             double* mean, double* var);

enum {
    GPL_VEC_FMEAN = 0, GPL_VEC_ALPHA = 1, GPL_VEC_Z = 2, GPL_VEC_V = 3, GPL_VEC_FVAR = 4,
    GPL_VEC_CAVITY_MEAN = 5, GPL_VEC_CAVITY_VAR = 6
};
/* gp$fit fields (R/laplace.R:119, R/ep.R:140-145). */
int gpl_model_vector(gpl_model* model, int32_t which, double* out /* n */);
/* C_chol = diag(sqrt V) chol(B)^T, n x n column-major lower (for gp_save). */
int gpl_model_chol(gpl_model* model, double* out);
/* gp_load (R/saveload.R:37-41) of a method_full fit saved by gp_save
 * (R/saveload.R:32-34): rebuilds a device model from the gp$fit fields
 * (R/laplace.R:119, R/ep.R:140-145) without refitting.  terms/x/ncols as for
 * gpl_fit (with the learned normalisation of the saved model); c_chol is the
 * saved n x n lower C_chol (as gpl_model_chol returns it; the part above the
 * diagonal is ignored); vectors[GPL_VEC_*] are the saved n-vectors (alpha and V
 * required, others may be NULL: zeros).  The restored model is FP64 whatever
 * precision it was fitted in, predicts, draws and (Gaussian) differentiates
 * like the original.  GPL_ERR_ARG for V <= 0; GPL_ERR_NOT_PD (minor = index)
 * for a non-positive diagonal of c_chol. */
int gpl_model_restore(gpl_ctx* ctx, const gpl_cf_term* terms, int32_t nterms, const double* x, int64_t n,
                      int64_t ncols, const gpl_lik* lik, const gpl_opts* opts, const double* c_chol,
                      const double* const* vectors, double log_evidence, int32_t iters, gpl_model** out);
int gpl_model_info(gpl_model* model, int64_t* n, int32_t* iters, double* log_evidence);
/* Precision bookkeeping of the last fit (new: the reference has one precision).
 * precision: GPL_PREC_MIXED, or GPL_PREC_FP64 for a plain fit and for a mixed fit
 * that was redone in FP64 (refit_fp64 = 1) because its iterative refinement did
 * not contract to a last relative correction of 1e-9 (cond(B) too large for the
 * FP32 factor) or the FP32 factor met a non-positive pivot; refine_sweeps and
 * refine_rel: sweeps and last relative correction of the mixed form's last
 * refinement (0 for a plain fit). */
int gpl_model_precision(gpl_model* model, int32_t* precision, int32_t* refit_fp64, int32_t* refine_sweeps,
                        double* refine_rel);
void gpl_model_free(gpl_model* model);

/* ---- FITC (method_fitc) ------------------------------------------------- */
/* Fully independent training conditional with m inducing inputs z:
 * prior covariance Q + diag(D), Q = K_xz Kz^-1 K_zx, Kz = K(z, z) + jitter I,
 * D = diag K(x, x) + jitter - diag Q (R/gp_fit.R:101-123).  Laplace Newton loop
 * for log-concave likelihoods on the inversion lemma (R/laplace.R:122-180,
 * 300-322, R/linear_algebra.R:3-172), one step for the Gaussian likelihood,
 * stopping when the unnormalised log posterior changes by less than opts->tol;
 * or damped parallel EP (approx EP: R/ep.R:148-185,332-370, tolerance 1e-3 as
 * there).  FP64 only.  z: m x ncols inducing
 * inputs prepared like x (column-major).  Status codes as gpl_fit. */
typedef struct gpl_fitc gpl_fitc;
int gpl_fitc_fit(gpl_ctx* ctx, const gpl_cf_term* terms, int32_t nterms, const double* x, int64_t n, int64_t ncols,
                 const double* z, int64_t m, const double* y, const double* trials, const double* offset,
                 const gpl_lik* lik, const gpl_opts* opts, gpl_fitc** out, double* log_evidence);
/* gp_pred_post.method_fitc (R/gp_pred.R:234-290, train = FALSE): mean
 * K_tz Kz^-1 K_zx alpha and variance prior_var - diag(A_t^T M A_t), with
 * A_t = Lz^-1 K_zt and M = A C^-1 A^T formed once after the fit. */
int gpl_fitc_pred(gpl_fitc* model, const double* xt, int64_t mt, int32_t want_var, double* mean, double* var);
/* gp_draw for method_fitc, marginal = FALSE (R/gp_draw.R:126-140): latent draws
 * mean + chol(cov + jitter I) z with the full FITC predictive covariance
 * A_t^T A_t + diag(Dt) - A_t^T M A_t (R/gp_pred.R:258-268); z and out as gpl_draw. */
int gpl_fitc_draw(gpl_fitc* model, const double* xt, int64_t mt, double jitter, const double* z, int64_t draws,
                  double* out);
/* Predictive variance at the training inputs with the FITC diagonal
 * (gp_pred_post.method_fitc with train = TRUE, R/gp_pred.R:270-285), as the
 * Laplace LOO uses it (R/gp_loo.R:120-130); out: n. */
int gpl_fitc_train_var(gpl_fitc* model, double* out);
/* fit vectors fmean, alpha, z, V, fvar and the EP cavity moments (GPL_VEC_*),
 * and the FITC diagonal D (which = 7) */
int gpl_fitc_vector(gpl_fitc* model, int32_t which, double* out);
int gpl_fitc_info(gpl_fitc* model, int64_t* n, int64_t* m, int32_t* iters, double* log_evidence);
void gpl_fitc_free(gpl_fitc* model);

/* ---- batched form: many independent Gaussian-likelihood GPs ------------- */
/* Energy and gradient of `nprob` independent GPs of uneven n in one device
 * pass (new: the reference fits one GP per R call).  All problems share the
 * covariance structure (term kinds, ops, n_lscale); inputs, n, d and
 * hyperparameters differ.  terms is nprob x nterms (row = problem).  Problems
 * are bucketed by padded size and each bucket runs the mixed-precision
 * pipeline (opts->refine_iters fixed refinement sweeps, default 4); the whole
 * pass is captured in one hipGraph at the first call and replayed. */
typedef struct gpl_batch gpl_batch;
int gpl_batch_create(gpl_ctx* ctx, int32_t nprob, const gpl_cf_term* terms, int32_t nterms,
                     const double* const* x, const int64_t* n, const int64_t* ncols,
                     const double* const* y, const double* sigma, const gpl_opts* opts,
                     gpl_batch** out);
/* New hyperparameters (same structure) for every problem. */
int gpl_batch_set_params(gpl_batch* batch, const gpl_cf_term* terms, const double* sigma);
/* log_evidence[nprob]; grad[nprob * gpl_batch_grad_len()] in the layout of gpl_grad_len.
 * A problem whose factor meets a non-positive pivot, or whose refinement does not
 * converge, gets log_evidence = -Inf and NaN gradients; the others are complete, and
 * the call returns GPL_ERR_NOT_PD (or, with no such problem, GPL_ERR_NOCONV) naming
 * the first such problem. */
int gpl_batch_energy_grad(gpl_batch* batch, double* log_evidence, double* grad);
int32_t gpl_batch_grad_len(gpl_batch* batch);
/* Number of buckets and the largest relative last refinement correction of the last pass. */
int gpl_batch_info(gpl_batch* batch, int32_t* nbuckets, double* refine_rel);
/* Per-problem outcome of the last gpl_batch_energy_grad: status[nprob] (GPL_OK,
 * GPL_ERR_NOT_PD, or GPL_ERR_NOCONV when the fixed refinement sweeps leave an
 * estimated error above 1e-8 of the solution -- last correction c2 times
 * rho / (1 - rho), rho = c2 / c1 the contraction of the last two sweeps; cond(B)
 * too large for the FP32 factor) and refine_rel[nprob] = c2; either may be NULL. */
int gpl_batch_status(gpl_batch* batch, int32_t* status, double* refine_rel);
/* Autotune: times GEMM tile edge {64,128} x outer panel {128,256,512,1024} per bucket,
 * keeps the fastest; writes up to cap rows (bucket n, tile, panel, ms) and the
 * number of buckets. */
int gpl_batch_tune(gpl_batch* batch, int32_t cap, int32_t* bucket_n, int32_t* bm, int32_t* nbo,
                   double* best_ms, int32_t* nbuckets);
/* Apply a persisted tuning table. */
int gpl_batch_set_tuning(gpl_batch* batch, int32_t count, const int32_t* bucket_n, const int32_t* bm,
                         const int32_t* nbo);
void gpl_batch_free(gpl_batch* batch);

/* gp_fit for many Gaussian-likelihood problems at once (new: the reference fits
 * one GP per call): the batched mixed-precision pass of gpl_batch_create (the
 * same covariance structure for all problems; x, n, ncols, y, sigma, opts as
 * there, opts NULL = 6 refinement sweeps) runs once, without a graph, and each
 * problem's factor, L^-1, alpha and fit vectors become a resident model, as
 * gpl_fit with GPL_PREC_MIXED leaves them (models[p] NULL: a new model; else an
 * existing model of ctx, refitted in place).  status[p] is GPL_OK, or
 * GPL_ERR_NOT_PD / GPL_ERR_NOCONV for a problem whose FP32 factor or refinement
 * failed (its models[p] is left as given: refit it with gpl_fit).  Returns
 * GPL_OK unless an argument or the device failed. */
int gpl_fit_batch(gpl_ctx* ctx, int32_t nprob, const gpl_cf_term* terms, int32_t nterms, const double* const* x,
                  const int64_t* n, const int64_t* ncols, const double* const* y, const double* sigma,
                  const gpl_opts* opts, gpl_model** models, double* log_evidence, int32_t* status);

/* ---- serving form: continuous batching of fit / predict requests -------- */
/* New (the reference answers one R call at a time).  Worker threads, each with
 * its own streams, take the oldest request that can run: predictions wait for
 * their model's fit, and every request waits until its estimated HBM footprint
 * fits the budget next to the resident models and the work in flight (idle
 * resident models are evicted least recently used first).  Queued predictions
 * against one model are served by one launch; every output is computed
 * independently, so results are the same alone and in a mix.  Inputs are
 * copied at submission; prediction outputs are written into the caller's
 * buffers, which must stay valid until gpl_server_wait returns. */
typedef struct gpl_server gpl_server;
int gpl_server_create(int32_t device, int32_t n_workers, int64_t mem_budget_bytes, gpl_server** out);
int gpl_server_submit_fit(gpl_server* srv, const gpl_cf_term* terms, int32_t nterms, const double* x,
                          int64_t n, int64_t ncols, const double* y, const double* trials,
                          const double* offset, const gpl_lik* lik, const gpl_opts* opts,
                          int64_t* model_id, int64_t* ticket);
int gpl_server_submit_pred(gpl_server* srv, int64_t model_id, const double* xt, int64_t m,
                           int32_t want_var, double* mean, double* var, int64_t* ticket);
/* Blocks until done; latency and queueing delay in ms, the fit's log evidence,
 * and how many requests the serving launch covered. */
int gpl_server_wait(gpl_server* srv, int64_t ticket, double* latency_ms, double* queue_ms,
                    double* log_evidence, int32_t* coalesced);
int gpl_server_release_model(gpl_server* srv, int64_t model_id);
int gpl_server_stats(gpl_server* srv, int64_t* fits, int64_t* preds, int64_t* launches,
                     int64_t* evictions, int64_t* bytes_in_use, int64_t* resident_models);
void gpl_server_destroy(gpl_server* srv);

/* Prediction on the streams of ctx (a model fitted on another context of the
 * same device can be used concurrently by several contexts). */
int gpl_pred_ctx(gpl_ctx* ctx, gpl_model* model, const double* xt, int64_t m, int32_t want_var,
                 double* mean, double* var);
/* gp_draw for method_full with marginal = FALSE (R/gp_draw.R:107-124,182-200):
 * latent draws mean + chol(cov + jitter I) z at the mt test points, where cov is
 * the full predictive covariance K_tt - aux^T aux, aux = C_chol^-1 K_t^T
 * (R/gp_pred.R:202-232, cov = TRUE).  z: mt x draws column-major standard normal
 * draws from the caller's generator (R's rnorm in mvnrnd); out: mt x draws.
 * GPL_ERR_NOT_PD when cov + jitter I has a non-positive pivot. */
int gpl_draw(gpl_model* model, const double* xt, int64_t mt, double jitter, const double* z, int64_t draws,
             double* out);
/* cfind (R/gp_pred.R:88-122 and R/gp_draw.R:3-35 with cfind; the covariance is
 * summed over the selected functions as in eval_cf.list, R/cf.R:402-411):
 * gpl_pred / gpl_draw with only the model's covariance terms terms[0..nsel)
 * (0-based indices into the gpl_cf_term array the model was fitted with, in
 * the order given; the first selected term starts a new sum).  Test inputs xt
 * are prepared as for the full model.  The posterior factor and alpha are the
 * fitted ones.  GPL_ERR_ARG for an index out of range or nsel outside [1, 8]. */
int gpl_pred_terms(gpl_model* model, const int32_t* terms, int32_t nsel, const double* xt, int64_t m,
                   int32_t want_var, double* mean, double* var);
int gpl_draw_terms(gpl_model* model, const int32_t* terms, int32_t nsel, const double* xt, int64_t mt,
                   double jitter, const double* z, int64_t draws, double* out);
/* gp_draw_prior.method_full (R/gp_draw.R:56-67): draws chol(K_tt + jitter I) z from
 * the prior of an unfitted model (gp_pred_prior with cov = TRUE, R/gp_pred.R:142-151;
 * mean 0).  terms / xt / ncols as for gpl_eval_cf, z and out as for gpl_draw. */
int gpl_draw_prior(gpl_ctx* ctx, const gpl_cf_term* terms, int32_t nterms, const double* xt, int64_t mt,
                   int64_t ncols, double jitter, const double* z, int64_t draws, double* out);
/* Device bytes held by a fitted model. */
int64_t gpl_model_bytes(gpl_model* model);
/* Frees the device buffers only the fit needed (K of a Laplace / EP fit, the mixed
 * form's FP64 copy of B and B^-1, workspaces); prediction, draws and gp_save read
 * what remains, and a later gradient or refit allocates again what it needs. */
int gpl_model_trim(gpl_model* model);

/* No counterpart in the reference: gpl_pred for callers that cannot allocate
 * device memory (serving loops that must not stall on hipMalloc / hipFree,
 * harnesses that forbid allocation).  Same arguments and results as gpl_pred
 * (bit for bit), with every scratch buffer taken from one static device array of
 * GPL_NOALLOC_SCRATCH_MB MiB per device that the library holds until the process
 * ends; the test points are processed in chunks that fit it.  A test point
 * needs 16 N bytes (N = n rounded up to 64) and a chunk at least 64 points, so
 * models up to n = 65408 are accepted; larger ones give GPL_ERR_ARG (use
 * gpl_pred).  Calls on one device are serialised by the library (they share the
 * array); the array keeps the last call's contents. */
#define GPL_NOALLOC_SCRATCH_MB 64
int gpl_pred_noalloc(gpl_model* model, const double* xt, int64_t m, int32_t want_var, double* mean, double* var);

#ifdef __cplusplus
}
#endif
#endif /* GPLITE_AMD_H */
