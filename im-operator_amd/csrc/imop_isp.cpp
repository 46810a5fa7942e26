// imop_isp.cpp -- C ABI of the RGB epilogue, lens shading and adaptive defect
// correction (include/imop_hip.h): parameter checks, the per-device gamma
// tables, the demosaic + epilogue entry point, the standalone in-place
// epilogue pass, the in-place lens-shading pass and the fixed-point passes of
// defect correction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
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

// Per-device state of defect correction: the scratch frame (input copy,
// pass control words, per-tile flags), grown as needed, and a pinned word
// block for reading the control words back.  One call at a time per device.
struct DefectState {
    std::mutex mu;
    void* d = nullptr;
    size_t bytes = 0;
    int* host = nullptr;
};
std::mutex g_defect_mu;
std::vector<std::unique_ptr<DefectState>> g_defect;

DefectState* defect_state(int dev) {
    std::lock_guard<std::mutex> lk(g_defect_mu);
    if ((int)g_defect.size() <= dev) g_defect.resize(dev + 1);
    if (!g_defect[dev]) g_defect[dev].reset(new DefectState);
    return g_defect[dev].get();
}

// Scratch of one call: the input copy, 256 bytes of pass control words and
// two flag bytes per tile.
struct DefectScratch {
    size_t frame, frame_al, need;
    DefectScratch(int w, int h, int bd) {
        frame = (size_t)w * (size_t)h * (bd > 8 ? 2u : 1u);
        frame_al = (frame + 255u) & ~(size_t)255u;
        need = frame_al + 256u + 2u * (size_t)defect_tiles_x(w) * (size_t)defect_tiles_y(h);
    }
};

// The replacement threshold: (int)(std::max(0.05f, std::min(1.0f, threshold))
// * max) (adaptive.cpp:17-18; NaN -> 1).
int defect_threshold(float threshold, int bd) {
    const float t_lo = (threshold < 1.0f) ? threshold : 1.0f;
    const float t = (0.05f < t_lo) ? t_lo : 0.05f;
    const int mv = bd <= 0 ? 255 : bd > 16 ? 65535 : (int)((1u << bd) - 1u);  // detail::safe_max_val
    return (int)(t * (float)mv);
}

// Stream-ordered form: the input copy, then `passes` passes (each returns at
// once after a pass that changed nothing), then the status word; the scratch
// is this call's own stream-ordered allocation (graph-owned under capture).
int defect_adaptive_async(uint8_t* data, int w, int h, int bd, float threshold, int passes, int* d_status,
                          hipStream_t st) {
    const DefectScratch sc(w, h, bd);
    void* d = nullptr;
    if (planes_alloc(sc.need, st, &d) != hipSuccess) return kInternal;
    uint8_t* orig = (uint8_t*)d;
    int* ctl = (int*)(orig + sc.frame_al);
    uint8_t* flags = orig + sc.frame_al + 256u;
    const int thresh = defect_threshold(threshold, bd);
    hipError_t e = launch_defect_copy(data, orig, sc.frame, ctl, st);
    for (int pass = 0; e == hipSuccess && pass < passes; ++pass)
        e = launch_defect_pass(data, orig, flags, ctl, w, h, bd, thresh, pass, st);
    if (e == hipSuccess) e = launch_defect_status(ctl, passes, d_status, st);
    const hipError_t fe = hipFreeAsync(d, st);
    return e == hipSuccess && fe == hipSuccess ? kOk : kInternal;
}

int defect_adaptive(uint8_t* data, int w, int h, int bd, float threshold, hipStream_t st) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return kInternal;
    DefectState* S = defect_state(dev);
    std::lock_guard<std::mutex> lk(S->mu);
    const DefectScratch sc(w, h, bd);
    const size_t frame = sc.frame, frame_al = sc.frame_al, need = sc.need;
    if (S->bytes < need) {
        if (S->d) {
            if (hipDeviceSynchronize() != hipSuccess) return kInternal;
            (void)hipFree(S->d);
            S->d = nullptr;
            S->bytes = 0;
        }
        if (hipMalloc(&S->d, need) != hipSuccess) return kInternal;
        S->bytes = need;
    }
    if (!S->host && hipHostMalloc((void**)&S->host, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
        S->host = nullptr;
        return kInternal;
    }
    uint8_t* orig = (uint8_t*)S->d;
    int* ctl = (int*)(orig + frame_al);
    uint8_t* flags = orig + frame_al + 256u;
    const int thresh = defect_threshold(threshold, bd);
    if (launch_defect_copy(data, orig, frame, ctl, st) != hipSuccess) return kInternal;
    // A pass fixes at least the next pixel of every chain of dependencies, and
    // the longest chain on a colour plane of pw x ph pixels is pw + 2 ph long.
    const long pw = (w - 4 + 1) / 2, ph = (h - 4 + 1) / 2;
    const long cap = pw + 2 * ph + 4;
    // Typical frames settle in two or three passes: check after four, then in
    // growing chunks (passes after convergence return at once).
    int pass = 0, chunk = 4;
    for (;;) {
        for (int i = 0; i < chunk; ++i, ++pass)
            if (launch_defect_pass(data, orig, flags, ctl, w, h, bd, thresh, pass, st) != hipSuccess)
                return kInternal;
        if (hipMemcpyAsync(S->host, ctl, 4 * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return kInternal;
        if (S->host[3]) return kInternal;
        if (S->host[(pass - 1) % 3] == 0) return kOk;
        if (pass >= cap) return kInternal;
        chunk = std::min(chunk * 2, 64);
    }
}

// defect_correct::validate_defect_correct_inputs (src/defect_correct/dispatch.cpp:7-15)
// and process_adaptive's size check (adaptive.cpp:13), then the buffer.
int defect_check(const void* d_data, uint64_t bytes, int width, int height, int bit_depth) {
    if (!d_data) return kNull;
    if (width <= 0 || height <= 0) return kDims;
    if (bit_depth < 1 || bit_depth > 16) return kBitDepth;
    if (width < 5 || height < 5) return kSmall;
    const uint64_t need = (uint64_t)width * (uint64_t)height * (bit_depth > 8 ? 2u : 1u);
    if (bytes < need || (bit_depth > 8 && ((uintptr_t)d_data & 1))) return kInternal;
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

int imop_hip_demosaic_pipeline(const void* d_in, uint64_t in_bytes, void* d_out, uint64_t out_bytes, int width,
                               int height, int pattern, int algorithm, int bit_depth, int is_packed,
                               const float* black_level, int black_level_mode,
                               const imop_hip_bayer_prologue* prologue, const imop_hip_rgb_epilogue* epilogue,
                               void* stream) {
    if (device_count_cached() == 0) return kInternal;
    int e = check_args(d_in, d_out, width, height, pattern, algorithm, bit_depth, is_packed, black_level,
                       black_level_mode);
    if (e != kOk) return e;
    Epilogue ep;
    e = make_epilogue(epilogue, bit_depth, &ep);
    if (e != kOk) return e;
    hipStream_t st = (hipStream_t)stream;
    const Epilogue* epp = ep.flags ? &ep : nullptr;
    if (!prologue || prologue->flags == 0)
        return run_demosaic(d_in, in_bytes, d_out, out_bytes, width, height, pattern, algorithm, bit_depth, is_packed,
                            black_level, black_level_mode, st, epp);
    if (prologue->flags & ~1) return kInternal;
    // run_demosaic's buffer checks
    if (in_bytes < bayer_bytes(width, height, bit_depth, is_packed) || out_bytes < rgb_bytes(width, height, bit_depth))
        return kInternal;
    if (bit_depth > 8 && ((!is_packed && ((uintptr_t)d_in & 1)) || ((uintptr_t)d_out & 1))) return kInternal;
    const LensParams lens = make_lens_params(width, height, pattern, &prologue->lens.coef[0][0],
                                             prologue->lens.center_x, prologue->lens.center_y);
    int mode = -1;
    const FrameParams f = make_params(d_in, in_bytes, d_out, width, height, pattern, bit_depth, is_packed,
                                      black_level, black_level_mode, algorithm, &mode);
    // the gain in the register-resident kernel's decode, for centres and
    // coefficients where its sequences are exact (its gamma table in LDS: at
    // most 4096 entries)
    const uint32_t lut_entries = f.out16 ? (uint32_t)f.maxv + 1u : 256u;
    if ((algorithm == kHqli || algorithm == kMg) && lens.fused_ok && fast_path_ok(f) &&
        (!(ep.flags & EPI_GAMMA) || lut_entries <= 4096u))
        return launch_fast(f, mode, st, epp, &lens) == hipSuccess ? kOk : kInternal;
    // else the prologue into a stream-ordered unpacked frame, demosaiced from there
    const uint64_t ub = bayer_bytes(width, height, bit_depth, 0);
    void* tmp = nullptr;
    if (planes_alloc(ub, st, &tmp) != hipSuccess) return kInternal;
    e = launch_bayer_prologue(f, lens, (uint8_t*)tmp, st) == hipSuccess ? kOk : kInternal;
    if (e == kOk)
        e = run_demosaic(tmp, ub, d_out, out_bytes, width, height, pattern, algorithm, bit_depth, 0, nullptr, 0, st,
                         epp);
    const hipError_t fe = hipFreeAsync(tmp, st);
    return e != kOk ? e : (fe == hipSuccess ? kOk : kInternal);
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

int imop_hip_defect_adaptive(void* d_data, uint64_t bytes, int width, int height, int pattern, int bit_depth,
                             float threshold, void* stream) {
    (void)pattern;  // same-colour neighbours at distance 2 whatever the CFA order
    if (device_count_cached() == 0) return kInternal;
    const int e = defect_check(d_data, bytes, width, height, bit_depth);
    if (e != kOk) return e;
    return defect_adaptive((uint8_t*)d_data, width, height, bit_depth, threshold, (hipStream_t)stream);
}

int imop_hip_defect_adaptive_async(void* d_data, uint64_t bytes, int width, int height, int pattern,
                                   int bit_depth, float threshold, int max_passes, int* d_status, void* stream) {
    (void)pattern;
    if (device_count_cached() == 0) return kInternal;
    const int e = defect_check(d_data, bytes, width, height, bit_depth);
    if (e != kOk) return e;
    if (!d_status) return kNull;
    if (max_passes < 1 || max_passes > 65536) return kInternal;
    return defect_adaptive_async((uint8_t*)d_data, width, height, bit_depth, threshold, max_passes, d_status,
                                 (hipStream_t)stream);
}

}  // extern "C"
