// Serving form: a continuous-batching request loop on one GPU.
//
// Fit and predict requests are queued as they arrive.  Worker threads (each
// with its own device context, i.e. its own HIP streams) take the oldest
// request that can run now: a prediction needs its model fitted, and every
// request needs its estimated HBM footprint to fit in the budget next to the
// resident models and the work in flight (idle resident models are evicted,
// least recently used first, to make room, but only for a request that is then
// admitted; an evicted model keeps its fit inputs on the host and is refitted
// when a prediction asks for it again).  Several workers keep requests
// overlapping on the device.  Queued predictions against the same resident
// model are coalesced into one launch (their test points are concatenated;
// every output column is computed independently of the others, so a result is
// the same alone and in a coalesced launch).  Fitted factors and alpha stay on
// the device and later predictions reuse them.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "serve_device.h"

using namespace gpl;


namespace {

enum { REQ_FIT = 0, REQ_PRED = 1 };
enum { ST_QUEUED = 0, ST_RUNNING = 1, ST_DONE = 2 };

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// The inputs of a fit, kept on the host while the model exists (an evicted model
// is refitted from them).
struct FitSpec {
    std::vector<gpl_cf_term> terms;
    std::vector<double> x, y, trials, offset;
    int64_t n = 0, ncols = 0;
    gpl_lik lik{};
    gpl_opts opts{};
};

struct Req {
    int64_t ticket = 0;
    int type = REQ_FIT;
    int64_t model_id = 0;
    bool internal = false;  // refit of an evicted model: nobody waits for it
    // fit
    std::shared_ptr<const FitSpec> spec;
    // prediction
    std::vector<double> xt;
    int64_t m = 0;
    int32_t want_var = 0;
    double* mean = nullptr;
    double* var = nullptr;
    // bookkeeping
    size_t footprint = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int state = ST_QUEUED;
    int status = GPL_OK;
    double log_evidence = 0;
    int32_t iters = 0;
    int32_t coalesced = 1;
};

struct Resident {
    gpl_model* model = nullptr;
    int64_t n = 0, ncols = 0;
    size_t bytes = 0;
    bool ready = false;      // fitted and resident
    bool fitting = false;    // a fit (first or after eviction) is queued or running
    bool evicted = false;    // fitted once, factors freed; refitted on the next prediction
    bool releasing = false;  // gpl_server_release_model is waiting for it
    bool failed = false;
    int status = GPL_OK;
    double last_use = 0;
    int refs = 0;  // queued or running predictions
    std::shared_ptr<const FitSpec> spec;
};

size_t fit_footprint(int64_t n, const gpl_lik& lik, const gpl_opts& o) {
    const size_t N = (size_t)serve_pad64(n), NN = N * N;
    const bool mixed = o.precision == GPL_PREC_MIXED;
    const bool gauss = lik.kind == GPL_LIK_GAUSSIAN;
    size_t b = NN * (mixed ? 4 : 8);                              // factor of B
    if (!gauss) b += NN * 8;                                      // stored K (FP64)
    if (o.approx == GPL_APPROX_EP) b += NN * 8;                   // L^-1 (EP)
    if (mixed) b += NN * 4 + NN * 8 + NN * 4;                     // L^-1, FP64 B, B^-1 (or the inverse's scratch)
    else if (o.flags & GPL_FLAG_INVERSE) b += 2 * NN * 8;         // L^-1, B^-1 (the inverse's scratch)
    const size_t T = N / 64;
    // matrix-vector partials: the larger of 32 chunks and two slots per four tile rows
    const size_t parts = std::max<size_t>(32, 2 * ((T + 3) / 4));
    return b + 2 * N * 64 * 8 + N * (kServeFeatCols + 40) * 8 + T * (T + 1) / 2 * 8 * 8 + N * 8 * parts + 4096;
}

size_t pred_footprint(int64_t n, int64_t m) {
    const size_t N = (size_t)serve_pad64(n), M = (size_t)serve_pad64(m);
    return 2 * N * M * 8 + M * (kServeFeatCols + 4) * 8;
}

}  // namespace

struct gpl_server {
    int device = 0;
    size_t budget = 0, in_use = 0;  // in_use: resident models + work in flight
    size_t inflight = 0;             // footprints of the jobs running now
    int64_t max_coalesce = 4096;
    int max_fit_batch = 16;  // Gaussian fits per batched pass
    std::vector<gpl_ctx*> ctxs;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::map<int64_t, std::unique_ptr<Req>> reqs;
    std::map<int64_t, Resident> models;
    // Device models no longer resident, kept for reuse by later fits (gpl_fit
    // recycles their buffers): freeing device memory waits for the whole device,
    // which would stall every worker.  Their bytes stay counted in in_use.
    struct Spare {
        gpl_model* model;
        size_t bytes;
        int worker;  // a model is refitted only on the context that made it
    };
    std::vector<Spare> spare;
    bool stop = false;
    int64_t next_ticket = 1, next_model = 1;
    int64_t done_fits = 0, done_preds = 0, launches = 0, evictions = 0;
};

static void release_bytes(gpl_server* s, size_t b) { s->in_use -= std::min(s->in_use, b); }

// Caller holds s->mu.  A model leaves residency: its device memory goes to the
// spare list (still counted) for a later fit on the same worker to reuse.
static void retire_model(gpl_server* s, gpl_model* m, size_t bytes, int worker) {
    if (m) s->spare.push_back({m, bytes, worker});
}

static bool evictable(const std::pair<const int64_t, Resident>& kv, int64_t keep) {
    const Resident& md = kv.second;
    return md.ready && md.refs == 0 && md.model && !md.releasing && kv.first != keep;
}

// Bytes that spare models and evictable idle resident models would give back.
static size_t reclaimable(gpl_server* s, int64_t keep) {
    size_t b = 0;
    for (auto& sp : s->spare) b += sp.bytes;
    for (auto& kv : s->models)
        if (evictable(kv, keep)) b += kv.second.bytes;
    return b;
}

// Caller holds s->mu.  Free spare models, then evict idle resident models (least
// recently used first), until `need` more bytes fit in the budget; `keep` is never
// evicted.  An evicted model keeps its entry and fit inputs.
static void evict_until(gpl_server* s, size_t need, int64_t keep) {
    while (s->in_use + need > s->budget && !s->spare.empty()) {
        gpl_model_free(s->spare.back().model);
        release_bytes(s, s->spare.back().bytes);
        s->spare.pop_back();
    }
    while (s->in_use + need > s->budget) {
        int64_t victim = -1;
        double oldest = 1e300;
        for (auto& kv : s->models)
            if (evictable(kv, keep) && kv.second.last_use < oldest) {
                oldest = kv.second.last_use;
                victim = kv.first;
            }
        if (victim < 0) return;
        Resident& md = s->models[victim];
        gpl_model_free(md.model);
        release_bytes(s, md.bytes);
        md.model = nullptr;
        md.bytes = 0;
        md.ready = false;
        md.evicted = true;
        s->evictions++;
    }
}

// Caller holds s->mu.  A spare device model of worker w to refit into (the
// smallest that holds `need` bytes, else the largest), or nullptr.  Its bytes
// leave in_use: the fit's footprint, counted from admission, covers them.
static gpl_model* take_spare(gpl_server* s, size_t need, int w) {
    int best = -1;
    for (int i = 0; i < (int)s->spare.size(); ++i) {
        if (s->spare[i].worker != w) continue;
        if (best < 0) {
            best = i;
            continue;
        }
        const size_t b = s->spare[i].bytes, bb = s->spare[best].bytes;
        const bool fits = b >= need, bfits = bb >= need;
        if ((fits && (!bfits || b < bb)) || (!fits && !bfits && b > bb)) best = i;
    }
    if (best < 0) return nullptr;
    gpl_model* m = s->spare[best].model;
    release_bytes(s, s->spare[best].bytes);
    s->spare.erase(s->spare.begin() + best);
    return m;
}

// A Gaussian mixed-precision fit without offset goes through gpl_fit_batch,
// together with the other such fits queued at the time (one batched pass).
static bool batchable(const FitSpec& f) {
    return f.lik.kind == GPL_LIK_GAUSSIAN && f.opts.precision == GPL_PREC_MIXED && f.opts.approx == GPL_APPROX_LAPLACE &&
           f.offset.empty();
}

// Same covariance structure and options: the fits can share a batched pass.
static bool same_structure(const FitSpec& a, const FitSpec& b) {
    if (a.terms.size() != b.terms.size() || a.ncols != b.ncols || a.opts.jitter != b.opts.jitter ||
        a.opts.refine_iters != b.opts.refine_iters)
        return false;
    for (size_t k = 0; k < a.terms.size(); ++k) {
        const gpl_cf_term &u = a.terms[k], &v = b.terms[k];
        if (u.kind != v.kind || (k > 0 && u.op != v.op) || u.col0 != v.col0 || u.ncols != v.ncols ||
            u.n_lscale != v.n_lscale || u.base_kind != v.base_kind)
            return false;
    }
    return true;
}

static void complete_with(gpl_server* s, Req* r, int status) {
    r->status = status;
    r->state = ST_DONE;
    r->t_start = r->t_done = now_ms();
    s->cv_done.notify_all();
}

// Caller holds s->mu.  Pick the oldest runnable request (and, for a prediction,
// the queued predictions on the same model that coalesce with it).  A request is
// admitted when its footprint fits the budget, after reclaiming spare and idle
// memory if that suffices; with nothing running it is admitted anyway, so the
// queue cannot stall.  Nothing is evicted for a request that is not admitted.
static std::vector<Req*> take_runnable(gpl_server* s) {
    std::vector<Req*> out;
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Req* r = *it;
        if (r->type == REQ_PRED) {
            auto mit = s->models.find(r->model_id);
            if (mit == s->models.end() || mit->second.failed) {  // unknown or failed model: complete with an error
                if (mit != s->models.end()) mit->second.refs--;
                it = s->queue.erase(it);
                complete_with(s, r, mit == s->models.end() ? GPL_ERR_ARG : mit->second.status);
                continue;
            }
            if (!mit->second.ready) {
                ++it;
                continue;
            }
        }
        const int64_t keep = r->type == REQ_PRED ? r->model_id : -1;
        if (s->in_use + r->footprint > s->budget) {
            const size_t rec = reclaimable(s, keep);
            const size_t floor = s->in_use - std::min(s->in_use, rec);
            if (floor + r->footprint > s->budget && s->inflight > 0) {
                ++it;
                continue;
            }
            evict_until(s, r->footprint, keep);
        }
        out.push_back(r);
        s->queue.erase(it);
        if (r->type == REQ_FIT && batchable(*r->spec)) {
            size_t fp = r->footprint;
            for (auto jt = s->queue.begin(); jt != s->queue.end() && out.size() < (size_t)s->max_fit_batch;) {
                Req* q = *jt;
                if (q->type == REQ_FIT && batchable(*q->spec) && same_structure(*r->spec, *q->spec) &&
                    s->in_use + fp + q->footprint <= s->budget) {
                    fp += q->footprint;
                    out.push_back(q);
                    jt = s->queue.erase(jt);
                } else {
                    ++jt;
                }
            }
        }
        if (r->type == REQ_PRED) {
            int64_t tot = r->m;
            size_t fp = r->footprint;
            for (auto jt = s->queue.begin(); jt != s->queue.end();) {
                Req* q = *jt;
                if (q->type == REQ_PRED && q->model_id == r->model_id && q->want_var == r->want_var &&
                    tot + q->m <= s->max_coalesce && s->in_use + fp + q->footprint <= s->budget) {
                    tot += q->m;
                    fp += q->footprint;
                    out.push_back(q);
                    jt = s->queue.erase(jt);
                } else {
                    ++jt;
                }
            }
        }
        return out;
    }
    return out;
}

static void worker_main(gpl_server* s, int w) {
    gpl_ctx* c = s->ctxs[w];
    for (;;) {
        std::vector<Req*> job;
        size_t fp = 0;
        gpl_model* m = nullptr;
        std::vector<gpl_model*> fm;  // fit jobs: one (recycled) model per request
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop) return;
                job = take_runnable(s);
                if (!job.empty()) break;
                s->cv_work.wait(lk);
            }
            const double t = now_ms();
            for (Req* r : job) {
                r->state = ST_RUNNING;
                r->t_start = t;
                r->coalesced = (int32_t)job.size();
                fp += r->footprint;
            }
            s->in_use += fp;
            s->inflight += fp;
            s->launches++;
            if (job[0]->type == REQ_FIT)
                for (Req* r : job) fm.push_back(take_spare(s, r->footprint, w));
            else
                m = s->models.find(job[0]->model_id)->second.model;  // refs > 0: resident
        }
        Req* r0 = job[0];
        TraceRange rg(r0->type == REQ_FIT ? "serve_fit_job" : "serve_pred_job");
        if (r0->type == REQ_FIT) {
            const size_t nj = job.size();
            std::vector<int> st(nj, GPL_OK);
            std::vector<double> lz(nj, 0.0);
            std::vector<char> done(nj, 0);
            if (batchable(*r0->spec)) {
                // one batched pass for every Gaussian fit taken (a single one included,
                // so a result does not depend on whether others arrived with it)
                std::vector<const double*> xs(nj), ys(nj);
                std::vector<int64_t> ns(nj), ncs(nj);
                std::vector<double> sig(nj);
                std::vector<gpl_cf_term> terms;
                for (size_t i = 0; i < nj; ++i) {
                    const FitSpec& f = *job[i]->spec;
                    xs[i] = f.x.data();
                    ys[i] = f.y.data();
                    ns[i] = f.n;
                    ncs[i] = f.ncols;
                    sig[i] = f.lik.sigma;
                    terms.insert(terms.end(), f.terms.begin(), f.terms.end());
                }
                std::vector<int32_t> bst(nj, GPL_OK);
                const int rb = gpl_fit_batch(c, (int32_t)nj, terms.data(), (int32_t)r0->spec->terms.size(), xs.data(),
                                             ns.data(), ncs.data(), ys.data(), sig.data(), &r0->spec->opts, fm.data(),
                                             lz.data(), bst.data());
                for (size_t i = 0; i < nj; ++i)
                    if (rb == GPL_OK && bst[i] == GPL_OK) done[i] = 1;
            }
            for (size_t i = 0; i < nj; ++i) {
                if (done[i]) continue;  // the others (also a batch member whose FP32 pass failed) fit alone
                const FitSpec& f = *job[i]->spec;
                st[i] = gpl_fit(c, f.terms.data(), (int32_t)f.terms.size(), f.x.data(), f.n, f.ncols, f.y.data(),
                                f.trials.empty() ? nullptr : f.trials.data(),
                                f.offset.empty() ? nullptr : f.offset.data(), &f.lik, &f.opts, &fm[i], &lz[i]);
            }
            // resident models keep what prediction reads: K, the FP64 copy of B and B^-1 go
            for (size_t i = 0; i < nj; ++i)
                if (fm[i] && (st[i] == GPL_OK || st[i] == GPL_ERR_NOCONV)) gpl_model_trim(fm[i]);
            std::unique_lock<std::mutex> lk(s->mu);
            release_bytes(s, fp);
            s->inflight -= std::min(s->inflight, fp);
            for (size_t i = 0; i < nj; ++i) {
                Req* r = job[i];
                gpl_model* mi = fm[i];
                r->status = st[i];
                r->log_evidence = lz[i];
                Resident& md = s->models.find(r->model_id)->second;  // release waits while fitting
                md.fitting = false;
                if (mi && (st[i] == GPL_OK || st[i] == GPL_ERR_NOCONV)) {
                    gpl_model_info(mi, nullptr, &r->iters, nullptr);
                    md.model = mi;
                    md.ready = true;
                    md.evicted = false;
                    md.bytes = (size_t)gpl_model_bytes(mi);
                    s->in_use += md.bytes;
                } else {
                    retire_model(s, mi, mi ? (size_t)gpl_model_bytes(mi) : 0, w);
                    if (mi) s->in_use += (size_t)gpl_model_bytes(mi);
                    md.failed = true;
                    md.status = st[i];
                }
                md.last_use = now_ms();
                r->state = ST_DONE;
                r->t_done = now_ms();
                s->done_fits++;
            }
            // the resident models may be larger than the admission estimates
            for (size_t i = 0; i < nj; ++i) evict_until(s, 0, job[i]->model_id);
            for (size_t i = 0; i < nj; ++i)
                if (job[i]->internal) s->reqs.erase(job[i]->ticket);
        } else {
            int64_t tot = 0;
            for (Req* r : job) tot += r->m;
            const int64_t ncols = (int64_t)(r0->xt.size() / std::max<int64_t>(r0->m, 1));
            std::vector<double> xt((size_t)tot * ncols), mean(tot), var(r0->want_var ? tot : 0);
            int64_t off = 0;
            for (Req* r : job) {  // column-major concatenation of the test inputs
                for (int64_t c2 = 0; c2 < ncols; ++c2)
                    memcpy(xt.data() + off + c2 * tot, r->xt.data() + c2 * r->m, r->m * 8);
                off += r->m;
            }
            int st = gpl_pred_ctx(c, m, xt.data(), tot, r0->want_var, mean.data(), r0->want_var ? var.data() : nullptr);
            off = 0;
            for (Req* r : job) {
                if (st == GPL_OK) {
                    memcpy(r->mean, mean.data() + off, r->m * 8);
                    if (r->want_var && r->var) memcpy(r->var, var.data() + off, r->m * 8);
                }
                off += r->m;
            }
            std::unique_lock<std::mutex> lk(s->mu);
            release_bytes(s, fp);
            s->inflight -= std::min(s->inflight, fp);
            Resident& md = s->models.find(r0->model_id)->second;
            md.refs -= (int)job.size();
            md.last_use = now_ms();
            const double t = now_ms();
            for (Req* r : job) {
                r->status = st;
                r->state = ST_DONE;
                r->t_done = t;
            }
            s->done_preds += (int64_t)job.size();
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

extern "C" int gpl_server_create(int32_t device, int32_t n_workers, int64_t mem_budget_bytes, gpl_server** out) {
    if (!out || n_workers < 1 || n_workers > 16 || mem_budget_bytes <= 0)
        return fail(GPL_ERR_ARG, "need 1..16 workers and a positive memory budget");
    gpl_server* s = new gpl_server();
    s->device = device;
    s->budget = (size_t)mem_budget_bytes;
    for (int w = 0; w < n_workers; ++w) {
        gpl_ctx* c = nullptr;
        int r = gpl_ctx_create(device, &c);
        if (r) {
            for (auto* cc : s->ctxs) gpl_ctx_destroy(cc);
            delete s;
            return r;
        }
        s->ctxs.push_back(c);
    }
    for (int w = 0; w < n_workers; ++w) s->workers.emplace_back(worker_main, s, w);
    *out = s;
    return GPL_OK;
}

extern "C" int gpl_server_submit_fit(gpl_server* s, const gpl_cf_term* terms, int32_t nterms, const double* x,
                                     int64_t n, int64_t ncols, const double* y, const double* trials,
                                     const double* offset, const gpl_lik* lik, const gpl_opts* opts,
                                     int64_t* model_id, int64_t* ticket) {
    if (!s || !terms || !x || !y || !lik || !model_id || !ticket || n < 1 || nterms < 1)
        return fail(GPL_ERR_ARG, "bad fit request");
    auto spec = std::make_shared<FitSpec>();
    spec->terms.assign(terms, terms + nterms);
    spec->x.assign(x, x + n * ncols);
    spec->y.assign(y, y + n);
    if (trials) spec->trials.assign(trials, trials + n);
    if (offset) spec->offset.assign(offset, offset + n);
    spec->n = n;
    spec->ncols = ncols;
    spec->lik = *lik;
    if (opts) spec->opts = *opts;
    else gpl_opts_default(&spec->opts);
    auto r = std::make_unique<Req>();
    r->type = REQ_FIT;
    r->footprint = fit_footprint(n, spec->lik, spec->opts);
    if (r->footprint > s->budget) return fail(GPL_ERR_OOM, "fit request exceeds the memory budget");
    r->spec = spec;
    r->t_submit = now_ms();
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next_ticket++;
    r->model_id = s->next_model++;
    Resident& md = s->models[r->model_id];
    md.n = n;
    md.ncols = ncols;
    md.fitting = true;
    md.spec = spec;
    *model_id = r->model_id;
    *ticket = r->ticket;
    s->queue.push_back(r.get());
    s->reqs[r->ticket] = std::move(r);
    s->cv_work.notify_one();
    return GPL_OK;
}

extern "C" int gpl_server_submit_pred(gpl_server* s, int64_t model_id, const double* xt, int64_t m,
                                      int32_t want_var, double* mean, double* var, int64_t* ticket) {
    if (!s || !xt || m < 1 || !mean || (want_var && !var) || !ticket) return fail(GPL_ERR_ARG, "bad pred request");
    auto r = std::make_unique<Req>();
    r->type = REQ_PRED;
    r->model_id = model_id;
    r->m = m;
    r->want_var = want_var;
    r->mean = mean;
    r->var = var;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->models.find(model_id);
    if (it == s->models.end() || it->second.releasing) return fail(GPL_ERR_ARG, "unknown model id");
    Resident& md = it->second;
    if (md.evicted && !md.fitting) {
        // refit the evicted model from its kept inputs; this prediction waits for it
        auto f = std::make_unique<Req>();
        f->type = REQ_FIT;
        f->internal = true;
        f->spec = md.spec;
        f->model_id = model_id;
        f->footprint = fit_footprint(md.spec->n, md.spec->lik, md.spec->opts);
        f->t_submit = now_ms();
        f->ticket = s->next_ticket++;
        md.fitting = true;
        md.evicted = false;
        s->queue.push_back(f.get());
        s->reqs[f->ticket] = std::move(f);
    }
    r->xt.assign(xt, xt + m * md.ncols);
    r->footprint = pred_footprint(it->second.n, m);
    r->t_submit = now_ms();
    r->ticket = s->next_ticket++;
    it->second.refs++;
    *ticket = r->ticket;
    s->queue.push_back(r.get());
    s->reqs[r->ticket] = std::move(r);
    s->cv_work.notify_one();
    return GPL_OK;
}

// Blocks until the request is done.  latency = completion - submission (ms);
// queue_ms = start - submission; coalesced = requests served by the same launch.
extern "C" int gpl_server_wait(gpl_server* s, int64_t ticket, double* latency_ms, double* queue_ms,
                               double* log_evidence, int32_t* coalesced) {
    if (!s) return fail(GPL_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(GPL_ERR_ARG, "unknown ticket");
    Req* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->state == ST_DONE; });
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    if (queue_ms) *queue_ms = r->t_start - r->t_submit;
    if (log_evidence) *log_evidence = r->log_evidence;
    if (coalesced) *coalesced = r->coalesced;
    const int st = r->status;
    s->reqs.erase(it);
    if (st != GPL_OK && st != GPL_ERR_NOCONV) return fail(st, "request failed with status " + std::to_string(st));
    return st;
}

extern "C" int gpl_server_release_model(gpl_server* s, int64_t model_id) {
    if (!s) return fail(GPL_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->models.find(model_id);
    if (it == s->models.end() || it->second.releasing) return fail(GPL_ERR_ARG, "unknown model id");
    it->second.releasing = true;  // no new predictions, never evicted, no second release
    // wait for queued and running predictions and any fit in progress; the entry
    // is looked up again after every wake (the map may have changed meanwhile)
    s->cv_done.wait(lk, [&] {
        const Resident& md = s->models.find(model_id)->second;
        return md.refs == 0 && !md.fitting;
    });
    it = s->models.find(model_id);
    if (it->second.model) {
        gpl_model_free(it->second.model);
        release_bytes(s, it->second.bytes);
    }
    s->models.erase(it);
    s->cv_work.notify_all();
    return GPL_OK;
}

extern "C" int gpl_server_stats(gpl_server* s, int64_t* fits, int64_t* preds, int64_t* launches,
                                int64_t* evictions, int64_t* bytes_in_use, int64_t* resident_models) {
    if (!s) return fail(GPL_ERR_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (fits) *fits = s->done_fits;
    if (preds) *preds = s->done_preds;
    if (launches) *launches = s->launches;
    if (evictions) *evictions = s->evictions;
    if (bytes_in_use) *bytes_in_use = (int64_t)s->in_use;
    if (resident_models) {
        int64_t k = 0;
        for (auto& kv : s->models)
            if (kv.second.model) ++k;
        *resident_models = k;
    }
    return GPL_OK;
}

extern "C" void gpl_server_destroy(gpl_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto& kv : s->models)
        if (kv.second.model) gpl_model_free(kv.second.model);
    for (auto& sp : s->spare) gpl_model_free(sp.model);
    for (auto* c : s->ctxs) gpl_ctx_destroy(c);
    delete s;
}
