// nh_debug.hip -- elementwise evaluation of device primitives (diagnostics).
#include "../../include/necsim_hip.h"
#include "nh_internal.h"
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
} // namespace

extern "C" int nh_debug_eval(nh_ctx *ctx, int32_t op, const uint64_t *in, uint64_t *out, size_t n) {
    if (!ctx || !in || !out || op < 0 || op > 11) return set_error(NH_E_INVALID, "nh_debug_eval: bad argument");
    if (n == 0) return NH_OK;
    const nh_ctx *c = ctx;
    size_t n_in = op == 7 || op == 8 ? 2 * n : (op >= 10 ? 3 * n : n);
    uint64_t *d_in = nullptr, *d_out = nullptr;
    if (hipSetDevice(c->device) != hipSuccess || hipMalloc(&d_in, n_in * 8) != hipSuccess ||
        hipMalloc(&d_out, n * 8) != hipSuccess) {
        if (d_in) (void)hipFree(d_in);
        return set_error(NH_E_NOMEM, "nh_debug_eval: allocation failed");
    }
    hipError_t e = hipMemcpy(d_in, in, n_in * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_debug_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, op, d_in, d_out, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d_out, n * 8, hipMemcpyDeviceToHost);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    if (e != hipSuccess) return set_error(NH_E_HIP, "nh_debug_eval: %s", hipGetErrorString(e));
    return NH_OK;
}
