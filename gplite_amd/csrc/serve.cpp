// Serving form: a continuous-batching request loop on one GPU.
//
// Fit and predict requests are queued as they arrive.  Worker threads (each
// with its own device context, i.e. its own HIP streams) take the oldest
// request that can run now: a prediction needs its model fitted, and every
// request needs its estimated HBM footprint to fit in the budget next to the
// resident models and the work in flight (idle resident models are evicted,
// least recently used first, to make room).  Several workers keep requests
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

#include "host_common.h"

using namespace gpl;

extern "C" void gpl_model_free(gpl_model* m);
extern "C" int64_t gpl_model_bytes(gpl_model* m);
extern "C" int gpl_model_info(gpl_model* m, int64_t* n, int32_t* iters, double* log_evidence);

namespace {

enum { REQ_FIT = 0, REQ_PRED = 1 };
enum { ST_QUEUED = 0, ST_RUNNING = 1, ST_DONE = 2 };

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
    int64_t ticket = 0;
    int type = REQ_FIT;
    int64_t model_id = 0;
    // fit
    std::vector<gpl_cf_term> terms;
    std::vector<double> x, y, trials, offset;
    int64_t n = 0, ncols = 0;
    gpl_lik lik{};
    gpl_opts opts{};
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
    bool ready = false;
    bool failed = false;
    int status = GPL_OK;
    double last_use = 0;
    int refs = 0;  // queued or running predictions
};

size_t fit_footprint(int64_t n, const gpl_lik& lik, const gpl_opts& o) {
    const size_t N = (size_t)pad64(n), NN = N * N;
    const bool mixed = o.precision == GPL_PREC_MIXED;
    const bool gauss = lik.kind == GPL_LIK_GAUSSIAN;
    size_t b = NN * (mixed ? 4 : 8);                              // factor of B
    if (!gauss) b += NN * 8;                                      // stored K (FP64)
    if (o.approx == GPL_APPROX_EP) b += NN * 8;                   // L^-1 (EP)
    if (mixed) b += NN * 4 + NN * 8 + (gauss ? NN * 4 : 0);       // L^-1, FP64 B, B^-1 (Gaussian)
    else if (o.flags & GPL_FLAG_INVERSE) b += NN * 8;             // L^-1
    const size_t T = N / 64;
    return b + 2 * N * 64 * 8 + N * (MAX_FEAT + 40) * 8 + T * (T + 1) / 2 * 8 * 8 + N * 8 * 32 + 4096;
}

size_t pred_footprint(int64_t n, int64_t m) {
    const size_t N = (size_t)pad64(n), M = (size_t)pad64(m);
    return 2 * N * M * 8 + M * (MAX_FEAT + 4) * 8;
}

}  // namespace

struct gpl_server {
    int device = 0;
    size_t budget = 0, in_use = 0;  // in_use: resident models + work in flight
    size_t inflight = 0;             // footprints of the jobs running now
    int64_t max_coalesce = 4096;
    std::vector<gpl_ctx*> ctxs;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::map<int64_t, std::unique_ptr<Req>> reqs;
    std::map<int64_t, Resident> models;
    bool stop = false;
    int64_t next_ticket = 1, next_model = 1;
    int64_t done_fits = 0, done_preds = 0, launches = 0, evictions = 0;
};

// Must hold s->mu.  Pick the oldest runnable request (and, for a prediction,
// the queued predictions on the same model that coalesce with it).
// Evict idle resident models (no queued or running predictions), least recently
// used first, until `need` more bytes fit in the budget; `keep` is never evicted.
// Caller holds s->mu.
static void evict_until(gpl_server* s, size_t need, int64_t keep) {
    while (s->in_use + need > s->budget) {
        int64_t victim = -1;
        double oldest = 1e300;
        for (auto& kv : s->models) {
            const Resident& md = kv.second;
            if (md.ready && md.refs == 0 && md.model && md.last_use < oldest && kv.first != keep) {
                oldest = md.last_use;
                victim = kv.first;
            }
        }
        if (victim < 0) return;
        Resident& md = s->models[victim];
        gpl_model_free(md.model);
        s->in_use -= std::min(s->in_use, md.bytes);
        s->models.erase(victim);
        s->evictions++;
    }
}

static std::vector<Req*> take_runnable(gpl_server* s) {
    std::vector<Req*> out;
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
        Req* r = *it;
        if (r->type == REQ_PRED) {
            auto mit = s->models.find(r->model_id);
            if (mit == s->models.end() || mit->second.failed) {  // unknown or failed model: complete with an error
                r->status = (mit == s->models.end()) ? GPL_ERR_ARG : mit->second.status;
                r->state = ST_DONE;
                r->t_start = r->t_done = now_ms();
                if (mit != s->models.end()) mit->second.refs--;
                s->queue.erase(it);
                s->cv_done.notify_all();
                return take_runnable(s);
            }
            if (!mit->second.ready) continue;
        }
        // admission by footprint, evicting idle resident models (LRU) if needed
        evict_until(s, r->footprint, r->type == REQ_PRED ? r->model_id : -1);
        // wait for memory while running work will free some; with nothing running and
        // nothing left to evict the request is admitted anyway, so the queue cannot stall
        if (s->in_use + r->footprint > s->budget && s->inflight > 0) continue;
        out.push_back(r);
        s->queue.erase(it);
        if (r->type == REQ_PRED) {
            int64_t tot = r->m;
            for (auto jt = s->queue.begin(); jt != s->queue.end();) {
                Req* q = *jt;
                if (q->type == REQ_PRED && q->model_id == r->model_id && q->want_var == r->want_var &&
                    tot + q->m <= s->max_coalesce && s->in_use + r->footprint + q->footprint <= s->budget) {
                    tot += q->m;
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
        }
        Req* r0 = job[0];
        if (r0->type == REQ_FIT) {
            gpl_model* m = nullptr;
            double lz = 0;
            int st = gpl_fit(c, r0->terms.data(), (int32_t)r0->terms.size(), r0->x.data(), r0->n, r0->ncols,
                             r0->y.data(), r0->trials.empty() ? nullptr : r0->trials.data(),
                             r0->offset.empty() ? nullptr : r0->offset.data(), &r0->lik, &r0->opts, &m, &lz);
            std::unique_lock<std::mutex> lk(s->mu);
            r0->status = st;
            r0->log_evidence = lz;
            Resident& md = s->models[r0->model_id];
            s->in_use -= std::min(s->in_use, fp);
            s->inflight -= std::min(s->inflight, fp);
            if (m && (st == GPL_OK || st == GPL_ERR_NOCONV)) {
                gpl_model_info(m, nullptr, &r0->iters, nullptr);
                md.model = m;
                md.ready = true;
                md.bytes = (size_t)gpl_model_bytes(m);
                s->in_use += md.bytes;
                md.last_use = now_ms();
                // the resident model may be larger than the admission estimate
                evict_until(s, 0, r0->model_id);
            } else {
                if (m) gpl_model_free(m);
                md.failed = true;
                md.status = st;
            }
            md.last_use = now_ms();
            r0->state = ST_DONE;
            r0->t_done = now_ms();
            s->done_fits++;
        } else {
            gpl_model* m;
            {
                std::unique_lock<std::mutex> lk(s->mu);
                m = s->models[r0->model_id].model;
            }
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
            int st = pred_on(c, m, xt.data(), tot, r0->want_var, mean.data(), r0->want_var ? var.data() : nullptr);
            off = 0;
            for (Req* r : job) {
                if (st == GPL_OK) {
                    memcpy(r->mean, mean.data() + off, r->m * 8);
                    if (r->want_var && r->var) memcpy(r->var, var.data() + off, r->m * 8);
                }
                off += r->m;
            }
            std::unique_lock<std::mutex> lk(s->mu);
            s->in_use -= std::min(s->in_use, fp);
            s->inflight -= std::min(s->inflight, fp);
            Resident& md = s->models[r0->model_id];
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
    auto r = std::make_unique<Req>();
    r->type = REQ_FIT;
    r->terms.assign(terms, terms + nterms);
    r->x.assign(x, x + n * ncols);
    r->y.assign(y, y + n);
    if (trials) r->trials.assign(trials, trials + n);
    if (offset) r->offset.assign(offset, offset + n);
    r->n = n;
    r->ncols = ncols;
    r->lik = *lik;
    if (opts) r->opts = *opts;
    else gpl_opts_default(&r->opts);
    r->footprint = fit_footprint(n, r->lik, r->opts);
    if (r->footprint > s->budget) return fail(GPL_ERR_OOM, "fit request exceeds the memory budget");
    r->t_submit = now_ms();
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next_ticket++;
    r->model_id = s->next_model++;
    Resident& md = s->models[r->model_id];
    md.n = n;
    md.ncols = ncols;
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
    if (it == s->models.end()) return fail(GPL_ERR_ARG, "unknown model id");
    r->xt.assign(xt, xt + m * it->second.ncols);
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
    if (it == s->models.end()) return fail(GPL_ERR_ARG, "unknown model id");
    s->cv_done.wait(lk, [&] { return it->second.refs == 0 && (it->second.ready || it->second.failed); });
    if (it->second.model) {
        gpl_model_free(it->second.model);
        s->in_use -= std::min(s->in_use, it->second.bytes);
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
    for (auto* c : s->ctxs) gpl_ctx_destroy(c);
    delete s;
}
