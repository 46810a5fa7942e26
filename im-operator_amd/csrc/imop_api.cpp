// imop_api.cpp -- the reference's GPU operator entry points, implemented on
// the gfx950 C ABI.  Linking this object in place of the reference's
// src/algorithms/cuda_dispatch.cpp and the black-level part of
// src/gpu_dispatch.cpp makes imop::demosaic() and
// black_level::process_black_level() take the HIP path unchanged.
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

// src/algorithms/cuda_dispatch.cpp:150-209: frames of one size, host buffers,
// results valid on return.
DemosaicError demosaic_cuda_batch(const uint8_t* const* bayer_data_array, uint8_t* const* rgb_data_array,
                                  int num_frames, int width, int height, BayerPattern pattern,
                                  DemosaicAlgorithm algorithm, int bit_depth) {
    if (!has_cuda()) return DemosaicError::InternalError;
    if (num_frames <= 0) return DemosaicError::Ok;
    if (!bayer_data_array || !rgb_data_array) return DemosaicError::NullInput;
    for (int i = 0; i < num_frames; ++i) {
        DemosaicError e = demosaic_cuda(bayer_data_array[i], rgb_data_array[i], width, height, pattern, algorithm,
                                        bit_depth, false);
        if (e != DemosaicError::Ok) return e;
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
