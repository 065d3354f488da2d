// Multithreaded exercise of the serving loop (serve.cpp) over the host-only
// stand-in of the device calls (test infrastructure, built with g++ under
// ThreadSanitizer or AddressSanitizer + UndefinedBehaviorSanitizer).  Client
// threads submit fits of every kind the loop routes differently (batched
// Gaussian, plain Gaussian, Laplace, fits that fail, batch members that are
// refitted alone), predictions against their own and older models, and
// releases while predictions are still queued; a small memory budget forces
// evictions and refits on demand.  Every prediction is checked against the
// stand-in's closed form of the same inputs.  Exit status 0 when all checks
// hold.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../../gplite_amd/csrc/serve_device.h"

double stub_logz(const double* x, int64_t n, int64_t ncols, const double* y, const gpl_lik* lik);
void stub_pred(double logz, const double* xt, int64_t m, int64_t ncols, double* mean, double* var);

namespace {

std::atomic<int> g_fail{0};
std::mutex g_print;

void check(bool ok, const char* what, int th, long a = 0) {
    if (ok) return;
    g_fail++;
    std::lock_guard<std::mutex> lk(g_print);
    std::fprintf(stderr, "thread %d: %s (%ld): %s\n", th, what, a, gpl_last_error());
}

struct Model {
    int64_t id = 0, n = 0, ncols = 0;
    double logz = 0;
    bool ok = true;
};

gpl_cf_term sexp_term(int ncols) {
    gpl_cf_term t{};
    t.kind = GPL_CF_SEXP;
    t.op = GPL_OP_SUM;
    t.ncols = ncols;
    t.n_lscale = 1;
    t.magn = 1.0;
    t.lscale[0] = 0.7;
    return t;
}

void client(gpl_server* s, int th, int rounds) {
    std::mt19937_64 rng(1000 + th);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<Model> mine;
    for (int it = 0; it < rounds; ++it) {
        // a fit of one of four kinds
        const int kind = (int)(rng() % 4);
        const int64_t ncols = 1 + (int64_t)(rng() % 3);
        int64_t n = 40 + (int64_t)(rng() % 300);
        if (kind == 0 && it % 3 == 0) n = 7 * (n / 7 + 1);  // its batched pass fails: refitted alone
        std::vector<double> x(n * ncols), y(n);
        for (auto& v : x) v = u(rng);
        for (auto& v : y) v = u(rng);
        const bool nan_fit = (rng() % 11) == 0;
        gpl_cf_term t = sexp_term((int)ncols);
        gpl_lik lik{kind == 3 ? GPL_LIK_BERNOULLI : GPL_LIK_GAUSSIAN, kind == 3 ? GPL_LINK_LOGIT : GPL_LINK_IDENTITY,
                    0.3, 0.0};
        if (kind == 3)
            for (auto& v : y) v = v > 0 ? 1.0 : 0.0;
        if (nan_fit) y[0] = NAN;
        gpl_opts o;
        gpl_opts_default(&o);
        o.precision = kind == 0 || kind == 1 ? GPL_PREC_MIXED : GPL_PREC_FP64;
        std::vector<double> offset;
        if (kind == 1) offset.assign(n, 0.0);  // mixed with an offset: not batched
        int64_t id = 0, tk = 0;
        int r = gpl_server_submit_fit(s, &t, 1, x.data(), n, ncols, y.data(), nullptr,
                                      offset.empty() ? nullptr : offset.data(), &lik, &o, &id, &tk);
        check(r == GPL_OK, "submit_fit", th, r);
        // the caller's arrays may go away once submitted
        const double lz = stub_logz(x.data(), n, ncols, y.data(), &lik);
        x.assign(x.size(), 9.0);
        y.assign(y.size(), 9.0);
        double le = 0;
        int32_t co = 0;
        // predictions submitted before the fit is waited for
        const int npred = 1 + (int)(rng() % 4);
        std::vector<std::vector<double>> xt(npred), mean(npred), var(npred);
        std::vector<int64_t> ptk(npred);
        std::vector<int64_t> pm(npred);
        for (int k = 0; k < npred; ++k) {
            pm[k] = 1 + (int64_t)(rng() % 200);
            xt[k].resize(pm[k] * ncols);
            for (auto& v : xt[k]) v = u(rng);
            mean[k].assign(pm[k], -7.0);
            var[k].assign(pm[k], -7.0);
            r = gpl_server_submit_pred(s, id, xt[k].data(), pm[k], k & 1, mean[k].data(), var[k].data(), &ptk[k]);
            check(r == GPL_OK, "submit_pred", th, r);
        }
        r = gpl_server_wait(s, tk, nullptr, nullptr, &le, &co);
        if (nan_fit) {
            check(r == GPL_ERR_NAN, "failed fit reports its status", th, r);
        } else {
            check(r == GPL_OK, "fit", th, r);
            check(le == lz, "log evidence", th, n);
        }
        for (int k = 0; k < npred; ++k) {
            r = gpl_server_wait(s, ptk[k], nullptr, nullptr, nullptr, &co);
            if (nan_fit) {
                check(r == GPL_ERR_NAN, "prediction on a failed model", th, r);
                continue;
            }
            check(r == GPL_OK, "prediction", th, r);
            std::vector<double> em(pm[k]), ev(pm[k]);
            stub_pred(lz, xt[k].data(), pm[k], ncols, em.data(), ev.data());
            for (int64_t i = 0; i < pm[k]; ++i) {
                check(mean[k][i] == em[i], "prediction mean", th, (long)i);
                if (k & 1) check(var[k][i] == ev[i], "prediction variance", th, (long)i);
            }
        }
        mine.push_back({id, n, ncols, lz, !nan_fit});
        // a prediction against an older model (evicted by now or not) and a release
        // while it may still be queued
        if (mine.size() > 2) {
            const size_t j = rng() % (mine.size() - 1);
            Model md = mine[j];
            const int64_t m = 1 + (int64_t)(rng() % 64);
            std::vector<double> xo(m * md.ncols), mo(m, -7.0), vo(m, -7.0);
            for (auto& v : xo) v = u(rng);
            int64_t t2 = 0;
            r = gpl_server_submit_pred(s, md.id, xo.data(), m, 1, mo.data(), vo.data(), &t2);
            check(r == GPL_OK, "submit_pred (older model)", th, r);
            const bool release_now = (rng() % 3) == 0;
            if (release_now) {
                r = gpl_server_release_model(s, md.id);
                check(r == GPL_OK, "release with a prediction queued", th, r);
                mine.erase(mine.begin() + j);
                int64_t t3 = 0;
                r = gpl_server_submit_pred(s, md.id, xo.data(), m, 1, mo.data(), vo.data(), &t3);
                check(r == GPL_ERR_ARG, "prediction on a released model is refused", th, r);
            }
            r = gpl_server_wait(s, t2, nullptr, nullptr, nullptr, nullptr);
            if (!md.ok) {
                check(r == GPL_ERR_NAN, "older failed model", th, r);
            } else {
                check(r == GPL_OK, "older prediction", th, r);
                std::vector<double> em(m), ev(m);
                stub_pred(md.logz, xo.data(), m, md.ncols, em.data(), ev.data());
                for (int64_t i = 0; i < m; ++i) check(mo[i] == em[i] && vo[i] == ev[i], "older model result", th, (long)i);
            }
        }
    }
    for (auto& md : mine) check(gpl_server_release_model(s, md.id) == GPL_OK, "final release", th);
}

}  // namespace

int main(int argc, char** argv) {
    const int nthreads = argc > 1 ? std::atoi(argv[1]) : 6;
    const int rounds = argc > 2 ? std::atoi(argv[2]) : 40;
    const int workers = argc > 3 ? std::atoi(argv[3]) : 4;
    // room for a few models of the sizes above: evictions and refits happen
    const int64_t budget = argc > 4 ? std::atoll(argv[4]) : 24ll << 20;
    gpl_server* s = nullptr;
    if (gpl_server_create(0, workers, budget, &s) != GPL_OK) {
        std::fprintf(stderr, "create: %s\n", gpl_last_error());
        return 2;
    }
    check(gpl_server_create(0, 0, budget, nullptr) == GPL_ERR_ARG, "bad create refused", -1);
    std::atomic<bool> done{false};
    std::thread watcher([&] {  // reads the statistics while requests run
        while (!done) {
            int64_t f, p, l, e, b, rm;
            check(gpl_server_stats(s, &f, &p, &l, &e, &b, &rm) == GPL_OK, "stats", -1);
            check(b >= 0 && rm >= 0, "stats values", -1);
            std::this_thread::sleep_for(std::chrono::milliseconds(2));
        }
    });
    std::vector<std::thread> cl;
    for (int t = 0; t < nthreads; ++t) cl.emplace_back(client, s, t, rounds);
    for (auto& t : cl) t.join();
    done = true;
    watcher.join();
    int64_t fits, preds, launches, evictions, bytes, resident;
    gpl_server_stats(s, &fits, &preds, &launches, &evictions, &bytes, &resident);
    std::printf("fits %ld preds %ld launches %ld evictions %ld bytes_in_use %ld resident %ld\n", (long)fits,
                (long)preds, (long)launches, (long)evictions, (long)bytes, (long)resident);
    check(resident == 0, "no model resident after every release", -1, (long)resident);
    check(fits >= (int64_t)nthreads * rounds, "every fit done", -1, (long)fits);
    check(launches < fits + preds, "some requests coalesced", -1, (long)launches);
    check(evictions > 0, "the budget forced evictions", -1, (long)evictions);
    gpl_server_destroy(s);
    std::printf("%s\n", g_fail ? "FAILED" : "OK");
    return g_fail ? 1 : 0;
}
