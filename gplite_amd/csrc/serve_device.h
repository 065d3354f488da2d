// What the serving loop (serve.cpp) needs from the rest of the library: the C
// ABI's fit / prediction / model calls, error reporting and a trace range.  It
// names no HIP type, so serve.cpp also builds with a host-only stand-in of these
// calls for sanitizer runs of its queueing, admission and eviction logic.
#pragma once

#include <stdint.h>

#include <string>

#include "../../include/gplite_amd.h"

namespace gpl {

int fail(int code, const std::string& msg);

// feature columns a model's device layout reserves per point (the MAX_FEAT of the kernels)
constexpr int kServeFeatCols = 64;

inline int64_t serve_pad64(int64_t n) { return (n + 63) / 64 * 64; }

// A named trace range (ROCTX) over one serving job.
class TraceRange {
public:
    explicit TraceRange(const char* name);
    ~TraceRange();
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace gpl
