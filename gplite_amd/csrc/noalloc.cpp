// Prediction without device allocation (gpl_pred_noalloc).
//
// gpl_pred stages the test inputs, their features, K_t^T and the triangular
// product through the context's scratch buffers, which it grows with hipMalloc
// (and hipFree, which waits for the whole device).  This variant takes that
// scratch from one file-scope device array of fixed size instead and walks the
// test points in chunks that fit it, so a serving loop, or a harness that
// forbids allocation, never allocates or frees device memory through it.  The
// kernels and their order are those of gpl_pred; every output column is
// computed independently of the others, so the results are gpl_pred's, bit for
// bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "host_common.h"

using namespace gpl;

namespace {

// The scratch array.  Its name is a host-side stand-in in host code: only the
// device address from hipGetSymbolAddress may be passed to kernels or copies.
constexpr size_t kScratchBytes = (size_t)GPL_NOALLOC_SCRATCH_MB << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;                      // guards g_ptr
void* g_ptr[kMaxDevices] = {};        // device address of `scratch` per device
std::mutex g_call_mu[kMaxDevices];    // one gpl_pred_noalloc at a time per device

}  // namespace

// Looks up the scratch array's device address on `device` (which the caller has
// made current) once; called by gpl_ctx_create.
int gpl::noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(GPL_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return GPL_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(GPL_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[device] = p;
    return GPL_OK;
}

extern "C" int gpl_pred_noalloc(gpl_model* m, const double* xt, int64_t mt, int32_t want_var, double* mean,
                                double* var) {
    if (!m || (!xt && mt > 0) || !mean || (want_var && !var)) return fail(GPL_ERR_ARG, "NULL argument");
    if (mt <= 0) return GPL_OK;
    gpl_ctx* c = model_ctx(m);
    const int dev = c->device;
    if (dev < 0 || dev >= kMaxDevices) return fail(GPL_ERR_ARG, "device index out of range");
    void* base;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = g_ptr[dev];
    }
    if (!base) return fail(GPL_ERR_HIP, "the no-allocation scratch is not set up on this device");
    int64_t N, ncols;
    model_dims(m, &N, &ncols);
    // largest chunk of test points (a multiple of 64) whose buffers fit the scratch
    int64_t M = 0;
    for (int64_t k = 64; k <= pad64(mt); k += 64) {
        if (pred_bufs_bytes(N, ncols, k) > kScratchBytes) break;
        M = k;
    }
    if (M == 0)
        return fail(GPL_ERR_ARG, "model too large for gpl_pred_noalloc: n = " + std::to_string(N) + " needs " +
                                     std::to_string(pred_bufs_bytes(N, ncols, 64) >> 20) + " MiB of scratch for 64 "
                                     "test points, the library holds " + std::to_string(GPL_NOALLOC_SCRATCH_MB) +
                                     " MiB (use gpl_pred)");
    PredBufs pb = pred_bufs_carve(base, N, ncols, M);
    pb.staged = true;
    std::lock_guard<std::mutex> lk(g_call_mu[dev]);
    for (int64_t i0 = 0; i0 < mt; i0 += M) {
        const int64_t mc = std::min(M, mt - i0);
        // rows i0 .. i0+mc of the column-major mt x ncols inputs into the staging
        // part of the scratch (leading dimension mc); pred_on waits for the stream
        // before returning, so the next chunk's copy cannot overtake this one
        if (ncols > 0) {
            std::lock_guard<std::mutex> ck(c->mu);
            HIP_TRY(hipSetDevice(dev));
            HIP_TRY(hipMemcpy2DAsync(pb.xt, mc * 8, xt + i0, mt * 8, mc * 8, ncols, hipMemcpyHostToDevice,
                                     c->stream));
        }
        int r = pred_on(c, m, xt + i0, mc, want_var, mean + i0, want_var ? var + i0 : nullptr, nullptr, 0, &pb);
        if (r) return r;
    }
    return GPL_OK;
}
