// imop_isp.cpp -- C ABI of the RGB epilogue and of lens shading
// (include/imop_hip.h): parameter checks, the per-device gamma tables, the
// demosaic + epilogue entry point, the standalone in-place epilogue pass and
// the in-place lens-shading pass.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "epilogue.hpp"
#include "host_common.hpp"
#include "imop_hip.h"
#include "plain.hpp"

namespace imop_hip {
namespace {

// detail::safe_max_val (src/tone/common.hpp:15-19)
int safe_max(int bd) { return bd <= 0 ? 255 : bd >= 16 ? 65535 : (1 << bd) - 1; }

// detail::clamp_val(static_cast<int>(v), mv) as the reference's x86-64 build
// computes it (cvttss2si: INT_MIN for NaN and out-of-range values)
int round_clamp_x86(float v, int mv) {
    const int i = (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : INT32_MIN;
    return i < 0 ? 0 : (i > mv ? mv : i);
}

// std::max(0.1f, std::min(10.0f, gamma)) (src/tone/gamma.cpp:17)
float gamma_clamp(float g) {
    const float lo = (g < 10.0f) ? g : 10.0f;
    return (0.1f < lo) ? lo : 0.1f;
}

// Gamma tables over every storable value (256 entries for u8 RGB, 65536 for
// u16), per device, bit depth and gamma; built once, never freed, so a
// captured graph can keep using one.
struct Lut {
    int dev, bd;
    uint32_t gamma_bits;
    uint16_t* d;
};
std::mutex g_lut_mu;
std::vector<Lut> g_luts;
constexpr size_t kMaxLuts = 64;

hipError_t gamma_lut(int bd, float gamma, const uint16_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    uint32_t bits;
    std::memcpy(&bits, &gamma, sizeof bits);
    std::lock_guard<std::mutex> lk(g_lut_mu);
    for (const Lut& l : g_luts)
        if (l.dev == dev && l.bd == bd && l.gamma_bits == bits) {
            *out = l.d;
            return hipSuccess;
        }
    if (g_luts.size() >= kMaxLuts) return hipErrorOutOfMemory;
    // tone::process_gamma (src/tone/gamma.cpp:18-31), one entry per value,
    // with the host's powf -- the function the reference calls
    const int mvi = safe_max(bd);
    const float mv = (float)mvi;
    const float inv_gamma = 1.0f / gamma_clamp(gamma);
    const int n = bd <= 8 ? 256 : 65536;
    std::vector<uint16_t> t(n);
    for (int v = 0; v < n; ++v) {
        const float norm = (float)v / mv;
        t[v] = (uint16_t)round_clamp_x86(fmaf(powf(norm, inv_gamma), mv, 0.5f), mvi);
    }
    void* d = nullptr;
    e = hipMalloc(&d, (size_t)n * sizeof(uint16_t));
    if (e != hipSuccess) return e;
    e = hipMemcpy(d, t.data(), (size_t)n * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return e;
    }
    g_luts.push_back(Lut{dev, bd, bits, (uint16_t*)d});
    *out = (const uint16_t*)d;
    return hipSuccess;
}

int make_epilogue(const imop_hip_rgb_epilogue* p, int bd, Epilogue* e) {
    std::memset(e, 0, sizeof *e);
    if (!p || p->flags == 0) return kOk;
    if (p->flags & ~7) return kInternal;
    e->flags = p->flags;
    e->maxv = safe_max(bd);
    for (int k = 0; k < 3; ++k) e->gain[k] = p->wb_gains[k];
    for (int k = 0; k < 9; ++k) e->m[k] = p->ccm[k];
    e->lut_max = bd <= 8 ? 255u : 65535u;
    // every stage value below 2^31 and never NaN: inputs are at most 65535,
    // so finite parameters with |gain| and each row's sum |m| at most 16384 do
    bool bounded = true;
    for (int k = 0; k < 3; ++k) bounded = bounded && std::isfinite(e->gain[k]) && std::fabs(e->gain[k]) <= 16384.0f;
    for (int r = 0; r < 3; ++r) {
        float s = 0.0f;
        for (int c = 0; c < 3; ++c) {
            bounded = bounded && std::isfinite(e->m[3 * r + c]);
            s += std::fabs(e->m[3 * r + c]);
        }
        bounded = bounded && s <= 16384.0f;
    }
    e->bounded = bounded ? 1 : 0;
    if (p->flags & EPI_GAMMA && gamma_lut(bd, p->gamma, &e->lut) != hipSuccess) return kInternal;
    return kOk;
}

}  // namespace
}  // namespace imop_hip

using namespace imop_hip;

extern "C" {

int imop_hip_demosaic_isp(const void* d_in, uint64_t in_bytes, void* d_out, uint64_t out_bytes, int width,
                          int height, int pattern, int algorithm, int bit_depth, int is_packed,
                          const float* black_level, int black_level_mode, const imop_hip_rgb_epilogue* epilogue,
                          void* stream) {
    if (device_count_cached() == 0) return kInternal;
    // the demosaic's own checks first, in the reference's order
    int e = check_args(d_in, d_out, width, height, pattern, algorithm, bit_depth, is_packed, black_level,
                       black_level_mode);
    if (e != kOk) return e;
    Epilogue ep;
    e = make_epilogue(epilogue, bit_depth, &ep);
    if (e != kOk) return e;
    return run_demosaic(d_in, in_bytes, d_out, out_bytes, width, height, pattern, algorithm, bit_depth, is_packed,
                        black_level, black_level_mode, (hipStream_t)stream, ep.flags ? &ep : nullptr);
}

int imop_hip_rgb_epilogue_apply(void* d_rgb, uint64_t rgb_bytes_, int width, int height, int bit_depth,
                                const imop_hip_rgb_epilogue* epilogue, void* stream) {
    if (device_count_cached() == 0) return kInternal;
    if (!d_rgb || !epilogue) return kNull;
    if (width <= 0 || height <= 0) return kDims;
    if (bit_depth < 1 || bit_depth > 16) return kBitDepth;
    if (rgb_bytes_ < rgb_bytes(width, height, bit_depth)) return kInternal;
    if (bit_depth > 8 && ((uintptr_t)d_rgb & 1)) return kInternal;
    Epilogue ep;
    const int e = make_epilogue(epilogue, bit_depth, &ep);
    if (e != kOk) return e;
    return launch_rgb_epilogue((uint8_t*)d_rgb, (uint64_t)width * (uint64_t)height, bit_depth > 8, ep,
                               (hipStream_t)stream) == hipSuccess
               ? kOk
               : kInternal;
}

int imop_hip_lens_shading_poly(void* d_data, uint64_t bytes, int width, int height, int pattern, int bit_depth,
                               const imop_hip_lens_shading* params, void* stream) {
    if (device_count_cached() == 0) return kInternal;
    // lens_shading::validate_lens_shading_inputs (src/lens_shading/dispatch.cpp:7-17)
    if (!d_data || !params) return kNull;
    if (width <= 0 || height <= 0) return kDims;
    if (bit_depth < 1 || bit_depth > 16) return kBitDepth;
    const uint64_t need = (uint64_t)width * (uint64_t)height * (bit_depth > 8 ? 2u : 1u);
    if (bytes < need || (bit_depth > 8 && ((uintptr_t)d_data & 1))) return kInternal;
    return launch_lens_shading((uint8_t*)d_data, width, height, pattern, bit_depth, &params->coef[0][0],
                               params->center_x, params->center_y, (hipStream_t)stream) == hipSuccess
               ? kOk
               : kInternal;
}

}  // extern "C"
