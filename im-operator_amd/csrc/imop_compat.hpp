// imop_compat.hpp -- declarations of the reference operator API that the
// gfx950 library implements, for building it without the reference tree.
//
// Names, enumerator order, struct layouts and signatures must equal the
// reference's (include/imop/types.hpp:12-39, include/imop/algorithms.hpp:
// 65-84; include/black_level/types.hpp, algorithms.hpp:40-48;
// include/white_balance/algorithms.hpp:58-62; include/ccm/algorithms.hpp:
// 61-65; include/tone/types.hpp:8-58, algorithms.hpp:46-51;
// include/lens_shading/types.hpp:8-60, algorithms.hpp:39-46;
// include/defect_correct/types.hpp:8-49, algorithms.hpp:39-46) so that the
// symbols defined in imop_api.cpp link against callers compiled with the
// reference headers (imop::demosaic, src/algorithms/dispatch.cpp:96-104, and
// the process_* dispatchers of black_level, white_balance, ccm, tone,
// lens_shading and defect_correct).  Define IMOP_WITH_REFERENCE_HEADERS to
// use the reference's own headers instead.
#pragma once

#include <cstdint>

#ifdef IMOP_WITH_REFERENCE_HEADERS
#include "imop/algorithms.hpp"
#include "black_level/algorithms.hpp"
#include "ccm/algorithms.hpp"
#include "defect_correct/algorithms.hpp"
#include "lens_shading/algorithms.hpp"
#include "tone/algorithms.hpp"
#include "white_balance/algorithms.hpp"
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

// The error enumerations of the RGB and Bayer stage modules share
// DemosaicError's numbering with InvalidChannels in place of InvalidPattern.
#define IMOP_STAGE_ERROR(Name) \
    enum class Name { Ok = 0, NullInput, InvalidDimensions, InvalidBitDepth, InvalidChannels, ImageTooSmall, InternalError }

namespace white_balance {
IMOP_STAGE_ERROR(WhiteBalanceError);
bool has_cuda();
WhiteBalanceError process_white_balance_cuda(const uint8_t* input, uint8_t* output, int width, int height,
                                             int channels, int bit_depth, float r_gain, float g_gain, float b_gain);
}  // namespace white_balance

namespace ccm {
IMOP_STAGE_ERROR(CCMError);
bool has_cuda();
CCMError process_ccm_cuda(const uint8_t* input, uint8_t* output, int width, int height, int channels, int bit_depth,
                          const float* matrix);
}  // namespace ccm

namespace tone {
enum class ToneAlgorithm { GAMMA, S_CURVE, LEVELS, CURVES_3POINT, SHADOWS_HIGHLIGHTS };
IMOP_STAGE_ERROR(ToneError);
struct ToneParams {
    float gamma = 1.0f;
    float contrast = 0.0f;
    float shadows = 0.0f;
    float highlights = 0.0f;
    float black_point = 0.0f;
    float white_point = 1.0f;
    float mid_point = 0.5f;
};
bool has_cuda();
ToneError process_tone_cuda(const uint8_t* input, uint8_t* output, int width, int height, int channels,
                            ToneAlgorithm algorithm, int bit_depth, const ToneParams& params);
}  // namespace tone

namespace lens_shading {
using imop::BayerPattern;
enum class LensShadingAlgorithm { POLYNOMIAL, FLAT_FIELD };
IMOP_STAGE_ERROR(LensShadingError);
struct ShadingPolynomial {
    float a2 = 0.0f;
    float a4 = 0.0f;
    float a6 = 0.0f;
};
struct LensShadingParams {
    ShadingPolynomial r_coef, gr_coef, gb_coef, b_coef;
    float center_x = 0.5f;
    float center_y = 0.5f;
    const uint8_t* flat_field = nullptr;
    int flat_field_width = 0;
    int flat_field_height = 0;
};
bool has_cuda();
LensShadingError process_lens_shading_cuda(uint8_t* data, int width, int height, BayerPattern pattern,
                                           LensShadingAlgorithm algorithm, int bit_depth,
                                           const LensShadingParams& params);
}  // namespace lens_shading

namespace defect_correct {
using imop::BayerPattern;
enum class DefectCorrectAlgorithm { ADAPTIVE, MAP_BASED };
IMOP_STAGE_ERROR(DefectCorrectError);
struct DefectPoint {
    int x, y;
};
struct DefectCorrectParams {
    float threshold = 0.3f;
    const DefectPoint* map = nullptr;
    int map_count = 0;
};
bool has_cuda();
DefectCorrectError process_defect_correct_cuda(uint8_t* data, int width, int height, BayerPattern pattern,
                                               DefectCorrectAlgorithm algorithm, int bit_depth,
                                               const DefectCorrectParams& params);
}  // namespace defect_correct

#undef IMOP_STAGE_ERROR
#endif
