// imop_chain.cpp -- C ABI of the differentiable ISP chain (include/imop_hip.h):
// argument checks, the per-device partial-sum buffer, launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "chain.hpp"
#include "host_common.hpp"
#include "imop_hip.h"

namespace imop_hip {
namespace {

// Per-workgroup partial sums: a fixed-size buffer per device (allocated on
// the device's first chain call, never resized, so graphs that capture a
// chain launch keep a valid pointer).  Forward and backward use disjoint
// halves; calls of one kind that overlap on one device race on theirs.
constexpr int kMaxBlocks = 2048;
constexpr size_t kFwdFloats = kMaxBlocks;
constexpr size_t kBwdFloats = (size_t)kMaxBlocks * 12;

std::mutex g_mu;
std::vector<float*> g_partials;

hipError_t partials_for_device(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_partials.size() <= dev) g_partials.resize(dev + 1, nullptr);
    if (!g_partials[dev]) {
        void* p = nullptr;
        e = hipMalloc(&p, (kFwdFloats + kBwdFloats) * sizeof(float));
        if (e != hipSuccess) return e;
        g_partials[dev] = (float*)p;
    }
    *out = g_partials[dev];
    return hipSuccess;
}

// Scratch of imop_hip_isp_chain_bwd_noalloc: the backward's per-workgroup
// partial gradient sums, kMaxBlocks x 12 floats = 96 KiB in a file-scope
// device array that every GPU the library runs on holds from the module's
// load until the process ends.  Host code never passes `scratch` itself (in
// host code the name is a stand-in, not a device address): the address of
// each device's copy is taken once per device with hipGetSymbolAddress and
// kept in g_noalloc_ptr.  Every call on one device shares that device's
// copy; each backward workgroup writes all 12 of its sums before the reduce
// reads them, so nothing needs clearing, and calls that overlap on one
// device race.
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float scratch[kBwdFloats];
std::mutex g_noalloc_mu;
float* g_noalloc_ptr[kMaxDevices] = {};

hipError_t noalloc_scratch(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_noalloc_ptr[dev] = (float*)p;
    }
    *out = g_noalloc_ptr[dev];
    return hipSuccess;
}

int check_chain(const float* raw, int n, int w, int h, int pattern, float max_value, const float* gains,
                const float* ccm, float gamma) {
    if (!raw || !gains || !ccm) return kNull;
    if (n <= 0 || w <= 0 || h <= 0) return kDims;
    if (pattern < 0 || pattern > 3) return kPattern;
    if (w < 6 || h < 6) return kSmall;  // the bilinear demosaic's minimum (hqli.cpp:11)
    if (!(max_value > 0.0f) || !(gamma > 0.0f)) return kInternal;
    if ((int64_t)n * w * h * 3 >= (int64_t)1 << 40) return kDims;
    return kOk;
}

ChainArgs make_args(const float* raw, int n, int w, int h, int pattern, float max_value, const float* gains,
                    const float* ccm, float gamma) {
    ChainArgs a{};
    a.raw = raw;
    a.n = n;
    a.w = w;
    a.h = h;
    a.pattern = pattern;
    a.inv_max = 1.0f / max_value;
    a.gains = gains;
    a.ccm = ccm;
    a.inv_gamma = 1.0f / gamma;
    a.mse_scale = (float)(2.0 / ((double)n * w * h * 3.0));
    return a;
}

}  // namespace
}  // namespace imop_hip

using namespace imop_hip;

extern "C" {

int imop_hip_isp_chain_fwd(const float* d_raw, int n, int width, int height, int pattern, float max_value,
                           const float* d_gains, const float* d_ccm, float gamma, float* d_out,
                           const float* d_target, float* d_loss, void* stream) {
    if (device_count_cached() == 0) return kInternal;
    int e = check_chain(d_raw, n, width, height, pattern, max_value, d_gains, d_ccm, gamma);
    if (e != kOk) return e;
    if (!d_out && !d_target) return kNull;
    if (d_target && !d_loss) return kNull;
    ChainArgs a = make_args(d_raw, n, width, height, pattern, max_value, d_gains, d_ccm, gamma);
    a.out = d_out;
    a.target = d_target;
    float* partials = nullptr;
    if (partials_for_device(&partials) != hipSuccess) return kInternal;
    a.partials = partials;
    const int64_t total = (int64_t)n * width * height;
    if ((width + 255) / 256 > kMaxBlocks) return kDims;
    const int blocks = chain_fwd_blocks(n, width, height, kMaxBlocks);
    hipStream_t st = (hipStream_t)stream;
    if (launch_chain_fwd(a, kMaxBlocks, st) != hipSuccess) return kInternal;
    if (d_target) {
        const float scale = (float)(1.0 / ((double)total * 3.0));
        if (launch_chain_reduce(partials, blocks, 1, scale, d_loss, st) != hipSuccess) return kInternal;
    }
    return kOk;
}

}  // extern "C"

namespace {

// The backward and the 12-value reduce with the per-workgroup sums in
// `partials` (at least kMaxBlocks x 12 floats).
int chain_bwd(const float* d_raw, int n, int width, int height, int pattern, float max_value, const float* d_gains,
              const float* d_ccm, float gamma, const float* d_grad_out, const float* d_target, float* d_grad_raw,
              float* d_grad_params, float* partials, hipStream_t st) {
    int e = check_chain(d_raw, n, width, height, pattern, max_value, d_gains, d_ccm, gamma);
    if (e != kOk) return e;
    if (!d_grad_out && !d_target) return kNull;
    if (!d_grad_raw && !d_grad_params) return kNull;
    ChainArgs a = make_args(d_raw, n, width, height, pattern, max_value, d_gains, d_ccm, gamma);
    a.grad_out = d_grad_out;
    a.target = d_grad_out ? nullptr : d_target;
    a.grad_raw = d_grad_raw;
    a.partials = partials;
    const int blocks = std::min(chain_bwd_tiles(n, width, height), kMaxBlocks);
    if (launch_chain_bwd(a, blocks, st) != hipSuccess) return kInternal;
    if (d_grad_params && launch_chain_reduce(a.partials, blocks, 12, 1.0f, d_grad_params, st) != hipSuccess)
        return kInternal;
    return kOk;
}

}  // namespace

extern "C" {

int imop_hip_isp_chain_bwd(const float* d_raw, int n, int width, int height, int pattern, float max_value,
                           const float* d_gains, const float* d_ccm, float gamma, const float* d_grad_out,
                           const float* d_target, float* d_grad_raw, float* d_grad_params, void* stream) {
    if (device_count_cached() == 0) return kInternal;
    int e = check_chain(d_raw, n, width, height, pattern, max_value, d_gains, d_ccm, gamma);
    if (e != kOk) return e;
    float* partials = nullptr;
    if (partials_for_device(&partials) != hipSuccess) return kInternal;
    return chain_bwd(d_raw, n, width, height, pattern, max_value, d_gains, d_ccm, gamma, d_grad_out, d_target,
                     d_grad_raw, d_grad_params, partials + kFwdFloats, (hipStream_t)stream);
}

int imop_hip_isp_chain_bwd_noalloc(const float* d_raw, int n, int width, int height, int pattern, float max_value,
                                   const float* d_gains, const float* d_ccm, float gamma, const float* d_grad_out,
                                   const float* d_target, float* d_grad_raw, float* d_grad_params, void* stream) {
    if (device_count_cached() == 0) return kInternal;
    int e = check_chain(d_raw, n, width, height, pattern, max_value, d_gains, d_ccm, gamma);
    if (e != kOk) return e;
    // The scratch holds kMaxBlocks workgroups' sums, the most the backward
    // ever launches (chain_bwd walks the tiles of any frame size with at most
    // that many), so every size that passes check_chain fits.
    static_assert(sizeof(scratch) == (size_t)kMaxBlocks * 12 * sizeof(float), "scratch holds kMaxBlocks x 12 sums");
    float* partials = nullptr;
    if (noalloc_scratch(&partials) != hipSuccess) return kInternal;
    return chain_bwd(d_raw, n, width, height, pattern, max_value, d_gains, d_ccm, gamma, d_grad_out, d_target,
                     d_grad_raw, d_grad_params, partials, (hipStream_t)stream);
}

int imop_hip_bilinear_fwd(const float* d_x, int width, int height, int pattern, float* d_out, void* stream) {
    if (device_count_cached() == 0) return kInternal;
    if (!d_x || !d_out) return kNull;
    if (width <= 0 || height <= 0) return kDims;
    if (pattern < 0 || pattern > 3) return kPattern;
    if (width < 6 || height < 6) return kSmall;
    return launch_bilinear_fwd(d_x, width, height, pattern, d_out, (hipStream_t)stream) == hipSuccess ? kOk
                                                                                                        : kInternal;
}

int imop_hip_bilinear_vjp(const float* d_v, int width, int height, int pattern, float* d_out, void* stream) {
    if (device_count_cached() == 0) return kInternal;
    if (!d_v || !d_out) return kNull;
    if (width <= 0 || height <= 0) return kDims;
    if (pattern < 0 || pattern > 3) return kPattern;
    if (width < 6 || height < 6) return kSmall;
    return launch_bilinear_vjp(d_v, width, height, pattern, d_out, (hipStream_t)stream) == hipSuccess ? kOk
                                                                                                       : kInternal;
}

}  // extern "C"
