// Host-only stand-in for the device calls the serving loop makes (test
// infrastructure): lets serve.cpp build with g++ under ThreadSanitizer,
// AddressSanitizer and UndefinedBehaviorSanitizer, so its queueing, admission,
// eviction, refit-on-demand and release logic runs with real threads.  Results
// are deterministic functions of the inputs (so a refit after an eviction gives
// the result of the first fit), and every call sleeps a little to let requests
// overlap.
#include <chrono>
#include <cmath>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../gplite_amd/csrc/serve_device.h"

struct gpl_ctx {
    int device = 0;
};

struct gpl_model {
    gpl_ctx* ctx = nullptr;
    int64_t n = 0, ncols = 0;
    double logz = 0;
    std::vector<double> alpha;  // touched by every prediction (heap use for ASan)
    int64_t bytes = 0;
    int32_t iters = 1;
};

static thread_local std::string g_err;

namespace gpl {
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
TraceRange::TraceRange(const char*) {}
TraceRange::~TraceRange() {}
}  // namespace gpl

extern "C" const char* gpl_last_error(void) { return g_err.c_str(); }

extern "C" int gpl_ctx_create(int32_t device, gpl_ctx** out) {
    if (!out || device != 0) return gpl::fail(GPL_ERR_ARG, "no such device");
    *out = new gpl_ctx();
    return GPL_OK;
}

extern "C" void gpl_ctx_destroy(gpl_ctx* c) { delete c; }

extern "C" void gpl_opts_default(gpl_opts* o) {
    *o = gpl_opts{};
    o->maxiter = 30;
    o->tol = 1e-4;
    o->jitter = 1e-6;
    o->refine_iters = 8;
}

// the stand-in's "log evidence": a fixed function of the inputs
double stub_logz(const double* x, int64_t n, int64_t ncols, const double* y, const gpl_lik* lik) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double xi = 0.0;
        for (int64_t j = 0; j < ncols; ++j) xi += x[i + j * n];
        s += y[i] * y[i] + 0.25 * xi;
    }
    return -0.5 * s - n * std::log(lik->kind == GPL_LIK_GAUSSIAN ? lik->sigma : 1.0);
}

// the stand-in's predictive mean and variance
void stub_pred(double logz, const double* xt, int64_t m, int64_t ncols, double* mean, double* var) {
    for (int64_t i = 0; i < m; ++i) {
        double s = 0.0;
        for (int64_t j = 0; j < ncols; ++j) s += xt[i + j * m];
        mean[i] = 1e-3 * logz * (1.0 + s);
        if (var) var[i] = 1.0 / (1.0 + std::fabs(mean[i]));
    }
}

static void work(int64_t units) { std::this_thread::sleep_for(std::chrono::microseconds(20 + units / 16)); }

static int fit_one(gpl_ctx* c, const double* x, int64_t n, int64_t ncols, const double* y, const gpl_lik* lik,
                   gpl_model** inout, double* log_evidence) {
    if (!c || !x || !y || !lik || !inout || n < 1) return gpl::fail(GPL_ERR_ARG, "bad fit");
    if (std::isnan(y[0])) return gpl::fail(GPL_ERR_NAN, "NaN in the targets");
    gpl_model* m = *inout;
    if (m && m->ctx != c) return gpl::fail(GPL_ERR_ARG, "the model to refit belongs to another context");
    if (!m) {
        m = new gpl_model();
        m->ctx = c;
    }
    work(n);
    m->n = n;
    m->ncols = ncols;
    m->logz = stub_logz(x, n, ncols, y, lik);
    m->alpha.assign(y, y + n);
    const int64_t N = (n + 63) / 64 * 64;
    m->bytes = 8 * N * N + 1024 * N;
    m->iters = lik->kind == GPL_LIK_GAUSSIAN ? 1 : 3;
    *inout = m;
    if (log_evidence) *log_evidence = m->logz;
    return GPL_OK;
}

extern "C" int gpl_fit(gpl_ctx* c, const gpl_cf_term* terms, int32_t nterms, const double* x, int64_t n,
                       int64_t ncols, const double* y, const double* trials, const double* offset, const gpl_lik* lik,
                       const gpl_opts* opts, gpl_model** inout, double* log_evidence) {
    (void)terms, (void)nterms, (void)trials, (void)offset, (void)opts;
    return fit_one(c, x, n, ncols, y, lik, inout, log_evidence);
}

extern "C" int gpl_fit_batch(gpl_ctx* c, int32_t nprob, const gpl_cf_term* terms, int32_t nterms,
                             const double* const* x, const int64_t* n, const int64_t* ncols, const double* const* y,
                             const double* sigma, const gpl_opts* opts, gpl_model** models, double* log_evidence,
                             int32_t* status) {
    (void)terms, (void)nterms, (void)opts;
    for (int p = 0; p < nprob; ++p) {
        if (n[p] % 7 == 0) {  // as a failed FP32 factor: the caller refits this one alone
            status[p] = GPL_ERR_NOT_PD;
            continue;
        }
        gpl_lik lik{GPL_LIK_GAUSSIAN, GPL_LINK_IDENTITY, sigma[p], 0.0};
        gpl_model* m = models[p];
        const int r = fit_one(c, x[p], n[p], ncols[p], y[p], &lik, &m, log_evidence ? log_evidence + p : nullptr);
        status[p] = r;
        if (r == GPL_OK) models[p] = m;
    }
    return GPL_OK;
}

extern "C" int gpl_pred_ctx(gpl_ctx* c, gpl_model* m, const double* xt, int64_t mt, int32_t want_var, double* mean,
                            double* var) {
    if (!c || !m || !xt || !mean || (want_var && !var)) return gpl::fail(GPL_ERR_ARG, "bad prediction");
    work(mt);
    volatile double touch = 0.0;
    for (double a : m->alpha) touch = touch + a;
    stub_pred(m->logz, xt, mt, m->ncols, mean, want_var ? var : nullptr);
    return GPL_OK;
}

extern "C" void gpl_model_free(gpl_model* m) { delete m; }

extern "C" int64_t gpl_model_bytes(gpl_model* m) { return m ? m->bytes : 0; }

extern "C" int gpl_model_info(gpl_model* m, int64_t* n, int32_t* iters, double* log_evidence) {
    if (!m) return gpl::fail(GPL_ERR_ARG, "NULL model");
    if (n) *n = m->n;
    if (iters) *iters = m->iters;
    if (log_evidence) *log_evidence = m->logz;
    return GPL_OK;
}

extern "C" int gpl_model_trim(gpl_model* m) {
    if (!m) return gpl::fail(GPL_ERR_ARG, "NULL model");
    m->bytes = m->bytes / 2;
    return GPL_OK;
}
