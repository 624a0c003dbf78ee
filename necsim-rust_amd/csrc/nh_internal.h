// nh_internal.h -- error reporting and device allocation helpers shared by
// the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/necsim_hip.h"

namespace nh {

// thread-local description of the calling thread's last failure
inline thread_local char g_last_error[512] = "";

inline int set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

// return NH_E_HIP with the failing call from a C-ABI entry point
#define NH_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return nh::set_error(NH_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                 __FILE__, __LINE__);                                         \
    } while (0)

// propagate a non-OK return code
#define NH_TRY(call)                \
    do {                            \
        int r_ = (call);            \
        if (r_ != NH_OK) return r_; \
    } while (0)

// hipMalloc of n elements (at least one); NH_E_NOMEM with a message on failure
template <class T> int dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess)
        return set_error(NH_E_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                         hipGetErrorString(e));
    return NH_OK;
}
template <class T> void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

} // namespace nh

using nh::dalloc;
using nh::dfree;
using nh::set_error;

// a device context: one device, one non-blocking stream (opaque in the C ABI)
struct nh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t cus = 0;
    std::string name;
    // device address of the library's file-scope scratch array on this
    // context's device (nh_sort_events_noalloc), resolved by nh_init
    void *scratch = nullptr;
};
