// imop_compat.hpp -- declarations of the reference operator API that the
// gfx950 library implements, for building it without the reference tree.
//
// Names, enumerator order and signatures must equal the reference's
// (include/imop/types.hpp:12-39, include/imop/algorithms.hpp:65-84,
// include/black_level/types.hpp, include/black_level/algorithms.hpp:40-48)
// so that the symbols defined in imop_api.cpp link against callers compiled
// with the reference headers (imop::demosaic, src/algorithms/dispatch.cpp:
// 96-104; black_level::process_black_level, src/black_level/dispatch.cpp:
// 53-58).  Define IMOP_WITH_REFERENCE_HEADERS to use the reference's own
// headers instead.
#pragma once

#include <cstdint>

#ifdef IMOP_WITH_REFERENCE_HEADERS
#include "imop/algorithms.hpp"
#include "black_level/algorithms.hpp"
#else
namespace imop {

enum class BayerPattern { RGGB, BGGR, GRBG, GBRG };
enum class DemosaicAlgorithm { SUPER_FAST, HQLI, MG, L7, DFPD, AHD, AMAZE, RCD, PRISM };
enum class DemosaicError { Ok = 0, NullInput, InvalidDimensions, InvalidBitDepth, InvalidPattern, ImageTooSmall, InternalError };

bool has_cuda();
const char* cuda_device_name();
void cuda_synchronize();
DemosaicError demosaic_cuda(const uint8_t* bayer_data, uint8_t* rgb_data, int width, int height, BayerPattern pattern,
                            DemosaicAlgorithm algorithm, int bit_depth = 8, bool is_packed = false);
DemosaicError demosaic_cuda_batch(const uint8_t* const* bayer_data_array, uint8_t* const* rgb_data_array,
                                  int num_frames, int width, int height, BayerPattern pattern,
                                  DemosaicAlgorithm algorithm, int bit_depth = 8);

}  // namespace imop

namespace black_level {

using imop::BayerPattern;

enum class BlackLevelAlgorithm { PER_CHANNEL, GLOBAL };
enum class BlackLevelError { Ok = 0, NullInput, InvalidDimensions, InvalidBitDepth, InvalidChannels, ImageTooSmall, InternalError };

struct BlackLevelParams {
    float r_offset = 0.0f;
    float gr_offset = 0.0f;
    float gb_offset = 0.0f;
    float b_offset = 0.0f;
};

bool has_cuda();
BlackLevelError process_black_level_cuda(uint8_t* data, int width, int height, BayerPattern pattern,
                                         BlackLevelAlgorithm algorithm, int bit_depth, const BlackLevelParams& params);

}  // namespace black_level
#endif
