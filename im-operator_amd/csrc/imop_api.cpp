// imop_api.cpp -- the reference's GPU operator entry points, implemented on
// the gfx950 C ABI.  Linking this library in place of the reference's
// src/algorithms/cuda_dispatch.cpp and src/gpu_dispatch.cpp makes
// imop::demosaic() and the process_* dispatchers of black_level,
// white_balance, ccm, tone, lens_shading and defect_correct take the HIP path
// unchanged (the other modules' GPU hooks come from the integration unit
// im-operator_amd/integration/modules_without_gpu.cpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "imop_compat.hpp"
#include "imop_hip.h"

namespace imop {

// src/algorithms/cuda_dispatch.cpp:26-33
bool has_cuda() { return imop_hip_device_count() > 0; }

// src/algorithms/cuda_dispatch.cpp:35-49
const char* cuda_device_name() { return imop_hip_device_name(); }

// src/algorithms/cuda_dispatch.cpp:211-213
void cuda_synchronize() { (void)imop_hip_synchronize(); }

// src/algorithms/cuda_dispatch.cpp:68-141.  Same host-buffer, synchronous
// contract; packed input is handled here instead of refused.
DemosaicError demosaic_cuda(const uint8_t* bayer_data, uint8_t* rgb_data, int width, int height, BayerPattern pattern,
                            DemosaicAlgorithm algorithm, int bit_depth, bool is_packed) {
    return static_cast<DemosaicError>(imop_hip_demosaic_host(bayer_data, rgb_data, width, height,
                                                             static_cast<int>(pattern), static_cast<int>(algorithm),
                                                             bit_depth, is_packed ? 1 : 0, nullptr, 0));
}

// src/algorithms/cuda_dispatch.cpp:150-209: unpacked frames of one size, host
// buffers, results valid on return.  The reference double-buffers one frame
// at a time through a captured graph; here the frames stream in chunks
// through pinned double-buffered staging, one batched launch per chunk
// (imop_hip_demosaic_batch_host).
DemosaicError demosaic_cuda_batch(const uint8_t* const* bayer_data_array, uint8_t* const* rgb_data_array,
                                  int num_frames, int width, int height, BayerPattern pattern,
                                  DemosaicAlgorithm algorithm, int bit_depth) {
    if (!has_cuda()) return DemosaicError::InternalError;
    if (num_frames <= 0) return DemosaicError::Ok;
    if (!bayer_data_array || !rgb_data_array) return DemosaicError::NullInput;
    for (int i = 0; i < num_frames; ++i)
        if (!bayer_data_array[i] || !rgb_data_array[i]) return DemosaicError::NullInput;
    if (width <= 0 || height <= 0) return DemosaicError::InvalidDimensions;
    if (bit_depth < 1 || bit_depth > 16) return DemosaicError::InvalidBitDepth;
    const size_t px = (size_t)width * (size_t)height;
    const size_t ib = px * (bit_depth <= 8 ? 1u : 2u), ob = px * 3u * (bit_depth > 8 ? 2u : 1u);
    std::vector<imop_hip_frame_desc> descs(num_frames, imop_hip_frame_desc{});
    for (int k = 0; k < num_frames; ++k) {
        imop_hip_frame_desc& d = descs[k];
        d.in = bayer_data_array[k];
        d.in_bytes = ib;
        d.out = rgb_data_array[k];
        d.out_bytes = ob;
        d.width = width;
        d.height = height;
        d.pattern = static_cast<int32_t>(pattern);
        d.algorithm = static_cast<int32_t>(algorithm);
        d.bit_depth = bit_depth;
    }
    return static_cast<DemosaicError>(imop_hip_demosaic_batch_host(descs.data(), num_frames, nullptr));
}

}  // namespace imop

namespace black_level {

bool has_cuda() { return imop_hip_device_count() > 0; }

// src/gpu_dispatch.cpp:421-448, with the CPU module's offset semantics.
BlackLevelError process_black_level_cuda(uint8_t* data, int width, int height, BayerPattern pattern,
                                         BlackLevelAlgorithm algorithm, int bit_depth, const BlackLevelParams& params) {
    const float offs[4] = {params.r_offset, params.gr_offset, params.gb_offset, params.b_offset};
    return static_cast<BlackLevelError>(imop_hip_black_level_host(data, width, height, static_cast<int>(pattern),
                                                                  static_cast<int>(algorithm), bit_depth, offs));
}

}  // namespace black_level

// ---------------------------------------------------------------------------
// The RGB and Bayer stage operators of src/gpu_dispatch.cpp.  Same host-buffer,
// synchronous contract as there: the caller's buffers are copied to a device
// buffer kept per host thread (the reference keeps a thread_local
// GpuWorkspace, gpu_dispatch.cpp:45), processed by the gfx950 kernels, and
// copied back before return.  The results are the reference's CPU functions'
// bytes (the CPU path the dispatchers fall back to), not the float
// normalised pipeline of its CUDA kernel.  Any non-Ok value makes the
// dispatcher run its CPU path.

namespace {

struct StageBuffer {
    void* d = nullptr;
    size_t cap = 0;
    int dev = -1;
    ~StageBuffer() {
        if (d) (void)hipFree(d);
    }
    void* get(size_t n) {
        int now = 0;
        if (hipGetDevice(&now) != hipSuccess) return nullptr;
        if (d && (now != dev || cap < n)) {
            (void)hipFree(d);
            d = nullptr;
            cap = 0;
        }
        if (!d) {
            if (hipMalloc(&d, n) != hipSuccess) {
                d = nullptr;
                return nullptr;
            }
            cap = n;
            dev = now;
        }
        return d;
    }
};
thread_local StageBuffer t_stage;

constexpr int kStageOk = 0, kStageNull = 1, kStageDims = 2, kStageBitDepth = 3, kStageChannels = 4,
              kStageInternal = 6;

// One RGB stage (or several, by flags) over an interleaved RGB image in host
// memory; in and out may be the same buffer.
int rgb_stage_host(const uint8_t* in, uint8_t* out, int w, int h, int ch, int bd, const imop_hip_rgb_epilogue& ep) {
    if (imop_hip_device_count() == 0) return kStageInternal;
    if (!in || !out) return kStageNull;
    if (w <= 0 || h <= 0) return kStageDims;
    if (bd < 1 || bd > 16) return kStageBitDepth;
    if (ch != 3) return kStageChannels;
    const size_t n = (size_t)w * (size_t)h * 3u * (bd > 8 ? 2u : 1u);
    void* d = t_stage.get(n);
    if (!d) return kStageInternal;
    if (hipMemcpy(d, in, n, hipMemcpyHostToDevice) != hipSuccess) return kStageInternal;
    const int e = imop_hip_rgb_epilogue_apply(d, n, w, h, bd, &ep, nullptr);
    if (e != 0) return e;
    return hipMemcpy(out, d, n, hipMemcpyDeviceToHost) == hipSuccess ? kStageOk : kStageInternal;
}

// An in-place pass over an unpacked Bayer frame in host memory.
template <typename Pass>
int bayer_stage_host(uint8_t* data, int w, int h, int bd, Pass pass) {
    if (imop_hip_device_count() == 0) return kStageInternal;
    if (!data) return kStageNull;
    if (w <= 0 || h <= 0) return kStageDims;
    if (bd < 1 || bd > 16) return kStageBitDepth;
    const size_t n = (size_t)w * (size_t)h * (bd > 8 ? 2u : 1u);
    void* d = t_stage.get(n);
    if (!d) return kStageInternal;
    if (hipMemcpy(d, data, n, hipMemcpyHostToDevice) != hipSuccess) return kStageInternal;
    const int e = pass(d, (uint64_t)n);
    if (e != 0) return e;
    return hipMemcpy(data, d, n, hipMemcpyDeviceToHost) == hipSuccess ? kStageOk : kStageInternal;
}

imop_hip_rgb_epilogue no_stages() {
    imop_hip_rgb_epilogue ep{};
    ep.wb_gains[0] = ep.wb_gains[1] = ep.wb_gains[2] = 1.0f;
    ep.ccm[0] = ep.ccm[4] = ep.ccm[8] = 1.0f;
    ep.gamma = 1.0f;
    return ep;
}

}  // namespace

namespace white_balance {

bool has_cuda() { return imop_hip_device_count() > 0; }

// src/gpu_dispatch.cpp:343-360 (MANUAL): the gains of process_manual_wb
// (src/white_balance/manual.cpp:8-41).
WhiteBalanceError process_white_balance_cuda(const uint8_t* input, uint8_t* output, int width, int height,
                                             int channels, int bit_depth, float r_gain, float g_gain, float b_gain) {
    imop_hip_rgb_epilogue ep = no_stages();
    ep.flags = 1;
    ep.wb_gains[0] = r_gain;
    ep.wb_gains[1] = g_gain;
    ep.wb_gains[2] = b_gain;
    return static_cast<WhiteBalanceError>(rgb_stage_host(input, output, width, height, channels, bit_depth, ep));
}

}  // namespace white_balance

namespace ccm {

bool has_cuda() { return imop_hip_device_count() > 0; }

// src/gpu_dispatch.cpp:319-336 (LINEAR_3X3): process_linear_3x3
// (src/ccm/linear_3x3.cpp:9-46), out = M * in with M row-major.
CCMError process_ccm_cuda(const uint8_t* input, uint8_t* output, int width, int height, int channels, int bit_depth,
                          const float* matrix) {
    if (!matrix) return CCMError::InternalError;
    imop_hip_rgb_epilogue ep = no_stages();
    ep.flags = 2;
    for (int k = 0; k < 9; ++k) ep.ccm[k] = matrix[k];
    return static_cast<CCMError>(rgb_stage_host(input, output, width, height, channels, bit_depth, ep));
}

}  // namespace ccm

namespace tone {

bool has_cuda() { return imop_hip_device_count() > 0; }

// src/gpu_dispatch.cpp:292-312: GAMMA runs process_gamma
// (src/tone/gamma.cpp:10-34, exponent clamped to [0.1, 10]); every other
// algorithm returns InternalError, so tone::process_tone runs its CPU path.
ToneError process_tone_cuda(const uint8_t* input, uint8_t* output, int width, int height, int channels,
                            ToneAlgorithm algorithm, int bit_depth, const ToneParams& params) {
    if (algorithm != ToneAlgorithm::GAMMA) return ToneError::InternalError;
    imop_hip_rgb_epilogue ep = no_stages();
    ep.flags = 4;
    ep.gamma = params.gamma;
    return static_cast<ToneError>(rgb_stage_host(input, output, width, height, channels, bit_depth, ep));
}

}  // namespace tone

namespace lens_shading {

bool has_cuda() { return imop_hip_device_count() > 0; }

// src/gpu_dispatch.cpp:457-486: POLYNOMIAL runs process_polynomial
// (src/lens_shading/polynomial.cpp:11-52); FLAT_FIELD returns InternalError,
// as there, so the dispatcher runs its CPU path.
LensShadingError process_lens_shading_cuda(uint8_t* data, int width, int height, BayerPattern pattern,
                                           LensShadingAlgorithm algorithm, int bit_depth,
                                           const LensShadingParams& params) {
    if (algorithm != LensShadingAlgorithm::POLYNOMIAL) return LensShadingError::InternalError;
    imop_hip_lens_shading p{};
    const ShadingPolynomial* c[4] = {&params.r_coef, &params.gr_coef, &params.gb_coef, &params.b_coef};
    for (int k = 0; k < 4; ++k) {
        p.coef[k][0] = c[k]->a2;
        p.coef[k][1] = c[k]->a4;
        p.coef[k][2] = c[k]->a6;
    }
    p.center_x = params.center_x;
    p.center_y = params.center_y;
    const int pat = static_cast<int>(pattern);
    return static_cast<LensShadingError>(bayer_stage_host(data, width, height, bit_depth, [&](void* d, uint64_t n) {
        return imop_hip_lens_shading_poly(d, n, width, height, pat, bit_depth, &p, nullptr);
    }));
}

}  // namespace lens_shading

namespace defect_correct {

bool has_cuda() { return imop_hip_device_count() > 0; }

// src/gpu_dispatch.cpp:529-571: ADAPTIVE runs process_adaptive
// (src/defect_correct/adaptive.cpp:8-61); MAP_BASED returns InternalError,
// so the dispatcher runs its CPU path.
DefectCorrectError process_defect_correct_cuda(uint8_t* data, int width, int height, BayerPattern pattern,
                                               DefectCorrectAlgorithm algorithm, int bit_depth,
                                               const DefectCorrectParams& params) {
    if (algorithm != DefectCorrectAlgorithm::ADAPTIVE) return DefectCorrectError::InternalError;
    const int pat = static_cast<int>(pattern);
    const float t = params.threshold;
    return static_cast<DefectCorrectError>(bayer_stage_host(data, width, height, bit_depth, [&](void* d, uint64_t n) {
        return imop_hip_defect_adaptive(d, n, width, height, pat, bit_depth, t, nullptr);
    }));
}

}  // namespace defect_correct
