// Host-side pieces shared by the C-ABI translation units: device context,
// per-kernel-class event profiling, device buffers and error reporting.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gpl_internal.h"

namespace gpl {

int fail(int code, const std::string& msg);
void set_minor(int64_t minor);

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::gpl::fail(GPL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

int prepare_prog(const gpl_cf_term* terms, int32_t nterms, int64_t ncols, DevProg& P);
// Blocked FP64 Cholesky of the N x N lower tiles of L in place (Dinv: the 64x64
// diagonal-block inverses, ldiag: log L_ii, info: first failing row + 1) and the
// triangular inverse X = L^-1 (all tiles written; T: optional N x N scratch,
// which selects the recursive-doubling form); potrf_blocked with X and Tm forms
// both, the inverse's front half beside the factor's last panels, and with rhs
// overwrites that vector (N rows) with L^-1 rhs
void potrf_blocked(gpl_ctx* c, double* L, int64_t N, double* Dinv, double* ldiag, int* info, double* X = nullptr,
                   double* Tm = nullptr, double* rhs = nullptr);
void trtri_blocked(gpl_ctx* c, const double* L, const double* Dinv, int64_t N, double* X, double* T = nullptr);
// Blocked triangular solve with one vector (trsv_lower_f64's arguments, b left
// unchanged): one launch on the context's solve flags when N / 64 <=
// kSolveMaxBlocks, else one launch per block.  info receives kPanelWaitTimeout
// if a device wait times out.
void trsv_lower(gpl_ctx* c, hipStream_t s, const double* L, int64_t ld, const double* Dinv, int64_t N, double* b,
                double* x, int trans, int* info);
// the error reported for an info word of kPanelWaitTimeout
int wait_timeout_fail();
// probabilists' Gauss-Hermite nodes and weights normalised to sum 1 (R/quadrature.R:1-44)
void gauss_hermite(int q, std::vector<double>& x, std::vector<double>& w);
// Device buffers of one prediction launch over M = pad64(mt) test points (see
// pred_bufs_bytes); when given to pred_on they replace the context's own scratch.
struct PredBufs {
    double *xt, *Ft, *KtT, *v, *A2;
    bool staged = false;  // the test inputs are already in xt (leading dimension mt)
};
// Bytes of the PredBufs for a model of padded size N with ncols input columns and
// M test-point rows, each part aligned to 256 bytes; carve them out of `base`.
size_t pred_bufs_bytes(int64_t N, int64_t ncols, int64_t M);
PredBufs pred_bufs_carve(void* base, int64_t N, int64_t ncols, int64_t M);
// padded size and input columns of a fitted model, and the context that owns it
void model_dims(const gpl_model* m, int64_t* N, int64_t* ncols);
gpl_ctx* model_ctx(const gpl_model* m);
// device address of gpl_pred_noalloc's scratch array on `device` (current), once
int noalloc_init(int device);
// terms/nsel (optional): predict with the model's terms terms[0..nsel) only (cfind);
// pb (optional): caller-provided device buffers instead of the context's scratch
int pred_on(gpl_ctx* c, gpl_model* m, const double* xt, int64_t mt, int32_t want_var, double* mean, double* var,
            const int32_t* terms = nullptr, int32_t nsel = 0, const PredBufs* pb = nullptr);

}  // namespace gpl

using gpl::fail;

inline int64_t pad64(int64_t n) { return (n + 63) / 64 * 64; }

// the likelihood's parameter the device kernels take: Gaussian sigma, beta-binomial phi
inline double lik_par(const gpl_lik& l) {
    return l.kind == GPL_LIK_GAUSSIAN ? l.sigma : (l.kind == GPL_LIK_BETABINOM ? l.phi : 0.0);
}

// ---------------------------------------------------------------------------
// context, profiling
// ---------------------------------------------------------------------------
struct ProfRec {
    int cls;
    hipEvent_t a, b;
    double flops, bytes;
};

struct ProfClass {
    std::string name;
    double ms = 0, flops = 0, bytes = 0;
    int64_t launches = 0;
};

struct gpl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // look-ahead panel work (default priority, like the main stream)
    hipStream_t aux = nullptr;   // the front half of L^-1 beside the Cholesky's last panels
    hipEvent_t ev_h = nullptr, ev_t = nullptr, ev_h2 = nullptr;
    hipEvent_t ev_a = nullptr, ev_p = nullptr, ev_d = nullptr;
    // flags and ticket counter of the single-launch panel factorisation
    // (kPanelSyncWords words, zero at creation), its launch epoch and tickets handed out so far,
    // and the polls after which one of its waits gives up (2^20; the
    // environment variable GPL_PANEL_WAIT_POLLS overrides it, which tests use to
    // reach the timeout path)
    unsigned* psync = nullptr;
    unsigned pepoch = 0, pbase = 0;
    // the same for the single-launch triangular solve (kSolveSyncWords words; its
    // launches take their epochs from pepoch too)
    unsigned* tsync = nullptr;
    unsigned tbase = 0;
    unsigned pwait = 1u << 20;
    bool prof = false;
    std::vector<ProfRec> pending;
    std::vector<hipEvent_t> pool;
    std::vector<ProfClass> classes;
    std::map<std::string, int> cls_index;
    std::vector<void*> scratch;
    std::vector<size_t> scratch_sz;
    std::mutex mu;

    hipEvent_t ev() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
    void resolve() {
        if (pending.empty()) return;
        hipEventSynchronize(pending.back().b);
        for (auto& r : pending) {
            float ms = 0;
            hipEventElapsedTime(&ms, r.a, r.b);
            auto& c = classes[r.cls];
            c.ms += ms;
            c.launches += 1;
            c.flops += r.flops;
            c.bytes += r.bytes;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    int cls(const char* name) {
        auto it = cls_index.find(name);
        if (it != cls_index.end()) return it->second;
        int id = (int)classes.size();
        classes.push_back(ProfClass{name});
        cls_index[name] = id;
        return id;
    }
    void* buf(int slot, size_t bytes) {
        if ((int)scratch.size() <= slot) {
            scratch.resize(slot + 1, nullptr);
            scratch_sz.resize(slot + 1, 0);
        }
        if (scratch_sz[slot] < bytes) {
            if (scratch[slot]) hipFree(scratch[slot]);
            scratch[slot] = nullptr;
            scratch_sz[slot] = 0;
            if (hipMalloc(&scratch[slot], bytes) != hipSuccess) return nullptr;
            scratch_sz[slot] = bytes;
        }
        return scratch[slot];
    }
};

// Column-major host matrix (rows x cols, leading dimension rows) to a device
// matrix of leading dimension dld, and back.  The host side goes through one
// contiguous copy into a staging buffer (context slot `slot`): a pageable 2D
// copy with short rows is carried out row by row.
inline hipError_t upload_2d(gpl_ctx* c, hipStream_t s, double* dst, int64_t dld, const double* src, int64_t rows,
                            int64_t cols, int slot) {
    double* st = (double*)c->buf(slot, (size_t)rows * cols * 8);
    if (!st) return hipErrorOutOfMemory;
    hipError_t e = hipMemcpyAsync(st, src, (size_t)rows * cols * 8, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipMemcpy2DAsync(dst, dld * 8, st, rows * 8, rows * 8, cols, hipMemcpyDeviceToDevice, s);
}

inline hipError_t download_2d(gpl_ctx* c, hipStream_t s, double* dst, const double* src, int64_t sld, int64_t rows,
                              int64_t cols, int slot) {
    double* st = (double*)c->buf(slot, (size_t)rows * cols * 8);
    if (!st) return hipErrorOutOfMemory;
    hipError_t e = hipMemcpy2DAsync(st, rows * 8, src, sld * 8, rows * 8, cols, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(dst, st, (size_t)rows * cols * 8, hipMemcpyDeviceToHost, s);
}

// A ROCTX range over one phase of a call (fit, gradient, prediction, batched
// pass, serving job), shown by rocprofv3 --marker-trace beside the kernels.
struct Range {
    explicit Range(const char* name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
    Range(const Range&) = delete;
    Range& operator=(const Range&) = delete;
};

// Times one launch group on the library stream when profiling is on.
struct Prof {
    gpl_ctx* c;
    ProfRec r;
    bool on;
    hipStream_t st;
    Prof(gpl_ctx* ctx, const char* name, double flops, double bytes, hipStream_t stream = nullptr)
        : c(ctx), on(ctx->prof), st(stream ? stream : ctx->stream) {
        if (!on) return;
        r.cls = c->cls(name);
        r.a = c->ev();
        r.b = c->ev();
        r.flops = flops;
        r.bytes = bytes;
        hipEventRecord(r.a, st);
    }
    ~Prof() {
        if (!on) return;
        hipEventRecord(r.b, st);
        c->pending.push_back(r);
        if (c->pending.size() > 8192) c->resolve();
    }
};

// ---------------------------------------------------------------------------
// model
// ---------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t sz = 0;
    void free_() {
        if (p) hipFree(p);
        p = nullptr;
        sz = 0;
    }
    bool ensure(size_t bytes) {
        if (sz >= bytes && p) return true;
        free_();
        if (bytes == 0) bytes = 8;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            return false;
        }
        sz = bytes;
        return true;
    }
    template <typename T>
    T* as() const { return reinterpret_cast<T*>(p); }
};

