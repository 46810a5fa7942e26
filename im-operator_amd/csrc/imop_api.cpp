// imop_api.cpp -- the reference's GPU operator entry points, implemented on
// the gfx950 C ABI.  Linking this object in place of the reference's
// src/algorithms/cuda_dispatch.cpp and the black-level part of
// src/gpu_dispatch.cpp makes imop::demosaic() and
// black_level::process_black_level() take the HIP path unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "imop_compat.hpp"
#include "imop_hip.h"

namespace {

// Per-host-thread device staging of demosaic_cuda_batch (grows, never shrinks).
struct BatchStaging {
    void* d = nullptr;
    size_t cap = 0;
    ~BatchStaging() {
        if (d) (void)hipFree(d);
    }
    bool ensure(size_t bytes) {
        if (cap >= bytes) return true;
        if (d) (void)hipFree(d);
        d = nullptr;
        cap = 0;
        if (hipMalloc(&d, bytes) != hipSuccess) return false;
        cap = bytes;
        return true;
    }
};
thread_local BatchStaging t_batch;

size_t align256(size_t v) { return (v + 255u) & ~(size_t)255u; }

}  // namespace

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
// at a time through a captured graph; here each chunk of frames (up to 1 GB
// staged) is copied in, demosaiced by one batched launch
// (imop_hip_demosaic_batch) and copied out.
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
    // 256-byte slots keep every frame on the aligned fast path
    const size_t si = align256(ib), so = align256(ob);
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)num_frames, ((size_t)1 << 30) / (si + so)));
    std::vector<imop_hip_frame_desc> descs;
    for (int start = 0; start < num_frames; start += chunk) {
        const int m = std::min(chunk, num_frames - start);
        if (!t_batch.ensure((size_t)m * (si + so))) return DemosaicError::InternalError;
        uint8_t* d_in = (uint8_t*)t_batch.d;
        uint8_t* d_out = d_in + (size_t)m * si;
        descs.assign(m, imop_hip_frame_desc{});
        for (int k = 0; k < m; ++k) {
            if (hipMemcpy(d_in + k * si, bayer_data_array[start + k], ib, hipMemcpyHostToDevice) != hipSuccess)
                return DemosaicError::InternalError;
            imop_hip_frame_desc& d = descs[k];
            d.in = d_in + k * si;
            d.in_bytes = ib;
            d.out = d_out + k * so;
            d.out_bytes = ob;
            d.width = width;
            d.height = height;
            d.pattern = static_cast<int32_t>(pattern);
            d.algorithm = static_cast<int32_t>(algorithm);
            d.bit_depth = bit_depth;
        }
        int bad = -1;
        const int e = imop_hip_demosaic_batch(descs.data(), m, nullptr, &bad);
        if (e != 0) return static_cast<DemosaicError>(e);
        for (int k = 0; k < m; ++k)
            if (hipMemcpy(rgb_data_array[start + k], d_out + k * so, ob, hipMemcpyDeviceToHost) != hipSuccess)
                return DemosaicError::InternalError;
    }
    return DemosaicError::Ok;
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
