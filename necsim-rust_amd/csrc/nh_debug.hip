// nh_debug.hip -- diagnostics: elementwise evaluation of device primitives
// and a plain copy kernel of known size (the HBM byte-counter calibration).
#include <cstring>
#include <vector>

#include "../../include/necsim_hip.h"
#include "nh_internal.h"
#include "nh_kernels.h"
#include "nh_rng.h"

using namespace nh;

namespace {
__global__ void k_debug_eval(int op, const uint64_t *in, uint64_t *out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = from_bits(in[i]);
    uint64_t r = 0;
    switch (op) {
    case 0: r = bits(ln(x)); break;
    case 1: r = bits(nh::exp(x)); break;
    case 2: r = bits(nh::sin(x)); break;
    case 3: r = bits(nh::cos(x)); break;
    case 4: r = bits(sqrt_(x)); break;
    case 5: r = bits(floor_(x)); break;
    case 6: r = bits(round_(x)); break;
    case 7: r = bits(from_bits(in[2 * i]) / from_bits(in[2 * i + 1])); break;
    case 8: r = wymum(in[2 * i], in[2 * i + 1]); break;
    case 9: r = seahash_diffuse(in[i]); break;
    case 10: r = prime(in[3 * i], in[3 * i + 1], in[3 * i + 2]); break;
    case 11: {
        uint64_t st = prime(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
        r = bits(uniform_closed_open(st));
        break;
    }
    default: r = 0;
    }
    out[i] = r;
}

// op 12: the Poisson event-time sampler's next event (poisson_next) of
// (seed, location, time) triples under one DevParams
__global__ void k_debug_poisson(DevParams P, const uint64_t *in, uint64_t *out, size_t m) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    DevParams Q = P;
    Q.seed = in[3 * i];
    const uint64_t lkey = prime_key(Q.seed, in[3 * i + 1]);
    uint64_t key = 0;
    const double et = poisson_next(Q, P.ratio, lkey, from_bits(in[3 * i + 2]), key);
    out[2 * i] = bits(et);
    out[2 * i + 1] = prime_keyed(lkey, key);
}
// 16-byte vector copy, grid-stride: reads and writes each byte exactly once
__global__ void __launch_bounds__(256) k_copy_calibrate(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                        size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_fill_calibrate(uint4 *dst, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = make_uint4((unsigned)i, (unsigned)(i >> 32), 0x9e3779b9u, 0x7f4a7c15u);
}
} // namespace

extern "C" int nh_copy_calibrate(nh_ctx *ctx, size_t bytes, int32_t reps, double *ms_per_copy) {
    if (!ctx || !ms_per_copy || reps < 1 || bytes < 16) return set_error(NH_E_INVALID, "nh_copy_calibrate: bad argument");
    const size_t n = bytes / 16;
    uint4 *src = nullptr, *dst = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess || hipMalloc(&src, n * 16) != hipSuccess ||
        hipMalloc(&dst, n * 16) != hipSuccess) {
        if (src) (void)hipFree(src);
        return set_error(NH_E_NOMEM, "nh_copy_calibrate: allocation failed");
    }
    hipStream_t st = ctx->stream;
    const unsigned blocks = (unsigned)(ctx->cus > 0 ? ctx->cus * 16 : 4096);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_calibrate, dim3(blocks), dim3(256), 0, st, src, n);
        hipLaunchKernelGGL(k_copy_calibrate, dim3(blocks), dim3(256), 0, st, src, dst, n); // warm-up
        e = hipEventRecord(e0, st);
    }
    for (int r = 0; e == hipSuccess && r < reps; r++) {
        hipLaunchKernelGGL(k_copy_calibrate, dim3(blocks), dim3(256), 0, st, src, dst, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e != hipSuccess) return set_error(NH_E_HIP, "nh_copy_calibrate: %s", hipGetErrorString(e));
    *ms_per_copy = (double)ms / reps;
    return NH_OK;
}

extern "C" int nh_debug_repeated_sum(double x, uint64_t n, double *out) {
    if (!out) return set_error(NH_E_INVALID, "nh_debug_repeated_sum: null output");
    *out = repeated_sum(x, n);
    return NH_OK;
}

extern "C" int nh_debug_presample(nh_ctx *ctx, double percentage, uint64_t n, int32_t serial, uint64_t *out) {
    if (!ctx || (!out && n)) return set_error(NH_E_INVALID, "nh_debug_presample: null argument");
    if (!(percentage > 0.0 && percentage < 1.0))
        return set_error(NH_E_INVALID, "the percentage pre-sampler takes 0 < p < 1 (got %g)", percentage);
    if (n == 0) return NH_OK;
    NH_HIP(hipSetDevice(ctx->device));
    return debug_presample(percentage, n, serial != 0, out, ctx->stream);
}

extern "C" int nh_debug_eval(nh_ctx *ctx, int32_t op, const uint64_t *in, uint64_t *out, size_t n) {
    if (!ctx || !in || !out || op < 0 || op > 12 || (op == 12 && n % 2))
        return set_error(NH_E_INVALID, "nh_debug_eval: bad argument");
    if (n == 0) return NH_OK;
    const nh_ctx *c = ctx;
    const size_t m = n / 2; // op 12: triples
    size_t n_in = op == 12 ? 3 * m + 2 : (op == 7 || op == 8 ? 2 * n : (op >= 10 ? 3 * n : n));
    uint64_t *d_in = nullptr, *d_out = nullptr;
    double *d_ratio = nullptr;
    if (hipSetDevice(c->device) != hipSuccess || hipMalloc(&d_in, n_in * 8) != hipSuccess ||
        hipMalloc(&d_out, n * 8) != hipSuccess ||
        (op == 12 && hipMalloc(&d_ratio, POISSON_TABLE_N * sizeof(double)) != hipSuccess)) {
        if (d_in) (void)hipFree(d_in);
        if (d_out) (void)hipFree(d_out);
        return set_error(NH_E_NOMEM, "nh_debug_eval: allocation failed");
    }
    hipError_t e = hipMemcpy(d_in, in, n_in * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && op == 12) {
        // the sampler's constants as a job evaluates them (poisson.rs:39-42)
        DevParams P{};
        const double lambda = from_bits(in[3 * m]), delta_t = from_bits(in[3 * m + 1]);
        P.lambda_dt = lambda * delta_t;
        P.p0 = nh::exp(-P.lambda_dt);
        P.lambda_dt_sqrt = sqrt_(P.lambda_dt);
        P.delta_t = delta_t;
        const uint64_t b = bits(delta_t);
        P.dt_pow2 = ((b >> 52) & 0x7ff) != 0 && (b & 0xfffffffffffffull) == 0;
        P.inv_delta_t = P.dt_pow2 ? 1.0 / delta_t : 0.0;
        std::vector<double> ratio(POISSON_TABLE_N, 0.0);
        P.prod_zero_at = poisson_table(P.lambda_dt, P.p0, ratio.data());
        e = hipMemcpy(d_ratio, ratio.data(), POISSON_TABLE_N * sizeof(double), hipMemcpyHostToDevice);
        P.ratio = d_ratio;
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_debug_poisson, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, 0, P, d_in, d_out, m);
            e = hipGetLastError();
        }
    } else if (e == hipSuccess) {
        hipLaunchKernelGGL(k_debug_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, op, d_in, d_out, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d_out, n * 8, hipMemcpyDeviceToHost);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    if (d_ratio) (void)hipFree(d_ratio);
    if (e != hipSuccess) return set_error(NH_E_HIP, "nh_debug_eval: %s", hipGetErrorString(e));
    return NH_OK;
}

// Sorts n host events by the PackedEvent order on the device, with the
// eager sort of the job and server paths (device_sized == 0) or the
// capturable one that reads its size from device memory (device_sized != 0).
template <class E>
static int debug_sort(nh_ctx *ctx, const E *in, size_t n, int32_t device_sized, E *out, const char *what) {
    if (!ctx || (n && (!in || !out))) return set_error(NH_E_INVALID, "%s: bad argument", what);
    if (n == 0) return NH_OK;
    if (n > 0xffffffffull) return set_error(NH_E_CAPACITY, "%s: at most 2^32 - 1 events", what);
    if (hipSetDevice(ctx->device) != hipSuccess) return set_error(NH_E_HIP, "%s: hipSetDevice", what);
    hipStream_t st = ctx->stream;
    SortBuffers sb{};
    E *d_in = nullptr, *d_out = nullptr;
    unsigned long long *d_range = nullptr;
    int rc = NH_OK;
    if ((rc = dalloc(&sb.keys[0], n)) || (rc = dalloc(&sb.keys[1], n)) || (rc = dalloc(&sb.vals[0], n)) ||
        (rc = dalloc(&sb.vals[1], n)) || (rc = dalloc(&sb.hist, sort_hist_words_staged(n))) ||
        (rc = dalloc(&sb.minmax, SORT_MINMAX_WORDS)) || (rc = dalloc(&d_in, n)) || (rc = dalloc(&d_out, n)) ||
        (rc = dalloc(&d_range, 2))) {
    } else {
        sb.capacity = n;
        sb.hist_words = sort_hist_words_staged(n);
        unsigned long long range[2] = {0ull, (unsigned long long)n}; // start, end
        hipError_t e = hipMemcpyAsync(d_in, in, n * sizeof(E), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(d_range, range, sizeof(range), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            if (device_sized) sort_events_device_sized(d_in, d_range + 1, d_range, n, d_out, sb, st);
            else sort_events(d_in, n, d_out, sb, st, false);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, n * sizeof(E), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = set_error(NH_E_HIP, "%s: %s", what, hipGetErrorString(e));
    }
    dfree(sb.keys[0]);
    dfree(sb.keys[1]);
    dfree(sb.vals[0]);
    dfree(sb.vals[1]);
    dfree(sb.hist);
    dfree(sb.minmax);
    dfree(d_in);
    dfree(d_out);
    dfree(d_range);
    return rc;
}

extern "C" int nh_debug_sort_events(nh_ctx *ctx, const nh_speciation_event *in, size_t n, int32_t device_sized,
                                    nh_speciation_event *out) {
    static_assert(sizeof(nh_speciation_event) == sizeof(DevEvent), "event record");
    return debug_sort(ctx, reinterpret_cast<const DevEvent *>(in), n, device_sized,
                      reinterpret_cast<DevEvent *>(out), "nh_debug_sort_events");
}

extern "C" int nh_debug_sort_packed_events(nh_ctx *ctx, const nh_packed_event *in, size_t n, int32_t device_sized,
                                           nh_packed_event *out) {
    static_assert(sizeof(nh_packed_event) == sizeof(DevPacked), "packed event record");
    return debug_sort(ctx, reinterpret_cast<const DevPacked *>(in), n, device_sized,
                      reinterpret_cast<DevPacked *>(out), "nh_debug_sort_packed_events");
}

// Sorts n host PackedEvents (in any order) and counts them as the round of a
// device_event_counter job does (the sort's permutation, the counts taken in
// its run windows or through the permutation), from a fresh reporter.
extern "C" int nh_debug_sort_count_packed_events(nh_ctx *ctx, const nh_packed_event *in, size_t n,
                                                 nh_event_counts *out) {
    const char *what = "nh_debug_sort_count_packed_events";
    if (!ctx || !out || (n && !in)) return set_error(NH_E_INVALID, "%s: bad argument", what);
    std::memset(out, 0, sizeof(*out));
    if (n == 0) return NH_OK;
    if (n > 0xffffffffull) return set_error(NH_E_CAPACITY, "%s: at most 2^32 - 1 events", what);
    if (hipSetDevice(ctx->device) != hipSuccess) return set_error(NH_E_HIP, "%s: hipSetDevice", what);
    hipStream_t st = ctx->stream;
    SortBuffers sb{};
    DevPacked *d_in = nullptr;
    EventCounts *d_c = nullptr;
    int rc = NH_OK;
    if ((rc = dalloc(&sb.keys[0], n)) || (rc = dalloc(&sb.keys[1], n)) || (rc = dalloc(&sb.vals[0], n)) ||
        (rc = dalloc(&sb.vals[1], n)) || (rc = dalloc(&sb.hist, sort_hist_words_staged(n))) ||
        (rc = dalloc(&sb.minmax, SORT_MINMAX_WORDS)) || (rc = dalloc(&d_in, n)) || (rc = dalloc(&d_c, 1))) {
    } else {
        sb.capacity = n;
        sb.hist_words = sort_hist_words_staged(n);
        EventCounts c{};
        hipError_t e = hipMemcpyAsync(d_in, in, n * sizeof(DevPacked), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(d_c, &c, sizeof(c), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            sort_and_count_packed(d_in, n, n, false, sb, d_c, st);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(&c, d_c, sizeof(c), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = set_error(NH_E_HIP, "%s: %s", what, hipGetErrorString(e));
        else std::memcpy(out, c.count, sizeof(*out));
    }
    dfree(sb.keys[0]);
    dfree(sb.keys[1]);
    dfree(sb.vals[0]);
    dfree(sb.vals[1]);
    dfree(sb.hist);
    dfree(sb.minmax);
    dfree(d_in);
    dfree(d_c);
    return rc;
}

// Counts n host PackedEvents (in PackedEvent order) with the device
// EventCounterReporter of device_event_counter jobs, from a fresh reporter.
extern "C" int nh_debug_count_packed_events(nh_ctx *ctx, const nh_packed_event *in, size_t n, nh_event_counts *out) {
    if (!ctx || !out || (n && !in)) return set_error(NH_E_INVALID, "nh_debug_count_packed_events: bad argument");
    std::memset(out, 0, sizeof(*out));
    if (n == 0) return NH_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_error(NH_E_HIP, "nh_debug_count_packed_events: hipSetDevice");
    hipStream_t st = ctx->stream;
    DevPacked *d_in = nullptr;
    EventCounts *d_c = nullptr;
    int rc = NH_OK;
    if ((rc = dalloc(&d_in, n)) || (rc = dalloc(&d_c, 1))) {
    } else {
        EventCounts c{};
        hipError_t e = hipMemcpyAsync(d_in, in, n * sizeof(DevPacked), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(d_c, &c, sizeof(c), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            launch_count_events(d_in, nullptr, nullptr, n, nullptr, d_c, st);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(&c, d_c, sizeof(c), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = set_error(NH_E_HIP, "nh_debug_count_packed_events: %s", hipGetErrorString(e));
        else std::memcpy(out, c.count, sizeof(*out));
    }
    dfree(d_in);
    dfree(d_c);
    return rc;
}
