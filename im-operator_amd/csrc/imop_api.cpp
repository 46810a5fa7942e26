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
