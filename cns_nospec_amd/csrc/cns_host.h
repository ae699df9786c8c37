// cns_host.h -- host-side helpers shared by the library's C ABI files (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "cns.h"

namespace cns {

// Stream of the calling thread's device work (cns_set_stream).
hipStream_t current_stream();

// Records the calling thread's error text (cns_last_error_string) and returns code.
int fail(int code, const char *fmt, ...);

}  // namespace cns

#define CNS_HIPCHK(call)                                                                           \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return ::cns::fail(CNS_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)
