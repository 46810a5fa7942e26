// modules_without_gpu.cpp -- integration unit for the reference's
// im_operator library when it links libimop_hip.so.
//
// The reference's src/gpu_dispatch.cpp defines the GPU hooks of every
// operator module (has_cuda() and process_<module>_cuda()), and its
// src/algorithms/cuda_dispatch.cpp those of the demosaic.  libimop_hip.so
// defines the hooks of imop (demosaic), black_level, white_balance, ccm,
// tone, lens_shading and defect_correct.  This unit, compiled in the
// reference's tree in place of those two files, gives the remaining
// modules -- which have no HIP implementation -- hooks that report no GPU,
// so their dispatchers take the CPU path they take on a build without a
// GPU: has_cuda() is false, and the process_*_cuda() functions (never
// reached while has_cuda() is false) return InternalError, the value on
// which every dispatcher falls back to the CPU.
//
// Signatures follow include/<module>/algorithms.hpp of the reference:
// color_temp :75-78, denoise :52-55, highlight_reconstruct :35-40,
// local_contrast :32-35, lut :89-92, saturation :44-47, sharpen :44-48.
#include "color_temp/algorithms.hpp"
#include "denoise/algorithms.hpp"
#include "highlight_reconstruct/algorithms.hpp"
#include "local_contrast/algorithms.hpp"
#include "lut/algorithms.hpp"
#include "saturation/algorithms.hpp"
#include "sharpen/algorithms.hpp"

namespace color_temp {
bool has_cuda() { return false; }
ColorTempError process_color_temp_cuda(const uint8_t*, uint8_t*, int, int, int, int, float, float, float) {
    return ColorTempError::InternalError;
}
}  // namespace color_temp

namespace denoise {
bool has_cuda() { return false; }
DenoiseError process_denoise_cuda(const uint8_t*, uint8_t*, int, int, int, DenoiseAlgorithm, int, float) {
    return DenoiseError::InternalError;
}
}  // namespace denoise

namespace highlight_reconstruct {
bool has_cuda() { return false; }
HighlightReconstructError process_highlight_reconstruct_cuda(const uint8_t*, uint8_t*, int, int, int,
                                                             HighlightReconstructAlgorithm, int,
                                                             const HighlightReconstructParams&) {
    return HighlightReconstructError::InternalError;
}
}  // namespace highlight_reconstruct

namespace local_contrast {
bool has_cuda() { return false; }
LocalContrastError process_local_contrast_cuda(const uint8_t*, uint8_t*, int, int, int, int,
                                               const LocalContrastParams&) {
    return LocalContrastError::InternalError;
}
}  // namespace local_contrast

namespace lut {
bool has_cuda() { return false; }
LUTError process_lut_cuda(const uint8_t*, uint8_t*, int, int, int, int, const LUT3D&) {
    return LUTError::InternalError;
}
}  // namespace lut

namespace saturation {
bool has_cuda() { return false; }
SaturationError process_saturation_cuda(const uint8_t*, uint8_t*, int, int, int, int, float, float) {
    return SaturationError::InternalError;
}
}  // namespace saturation

namespace sharpen {
bool has_cuda() { return false; }
SharpenError process_sharpen_cuda(const uint8_t*, uint8_t*, int, int, int, SharpenAlgorithm, int,
                                  const SharpenParams&) {
    return SharpenError::InternalError;
}
}  // namespace sharpen
