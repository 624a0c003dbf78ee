// nh_internal.h -- error reporting shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

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

} // namespace nh

using nh::set_error;

// a device context: one device, one non-blocking stream (opaque in the C ABI)
struct nh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t cus = 0;
    std::string name;
};
