// host_common.hpp -- host-side helpers shared by the C ABI translation units:
// argument checks with the reference's error codes, frame parameter setup,
// and the per-frame launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

struct imop_hip_frame_desc;

namespace imop_hip {

// imop::DemosaicError (include/imop/types.hpp:31-39)
enum : int { kOk = 0, kNull = 1, kDims = 2, kBitDepth = 3, kPattern = 4, kSmall = 5, kInternal = 6 };
// imop::DemosaicAlgorithm (include/imop/types.hpp:19-29)
enum : int { kSuperFast = 0, kHqli = 1, kMg = 2, kL7 = 3, kDfpd = 4, kAhd = 5, kAmaze = 6, kRcd = 7, kPrism = 8 };

// pixel::compute_bayer_byte_size / compute_rgb_byte_size (pixel_utils.hpp:21-34)
uint64_t bayer_bytes(int w, int h, int bd, int packed);
uint64_t rgb_bytes(int w, int h, int bd);

// Argument checks in imop::demosaic_cuda's order (cuda_dispatch.cpp:70-90).
int check_args(const void* in, const void* out, int w, int h, int pattern, int algo, int bd, int packed,
               const float* bl, int bl_mode);

// Black-level offsets by CFA parity (y & 1) * 2 + (x & 1), each max(0, .).
void bl_by_parity(int pattern, const float* offs, int mode, float q[4]);

// Kernel parameters of one frame; *mode = the HQLI / MG kernel mode (-1 for PRISM).
FrameParams make_params(const void* in, uint64_t in_bytes, void* out, int w, int h, int pattern, int bd, int packed,
                        const float* bl, int bl_mode, int algo, int* mode);

// Enqueue one frame: PRISM / RCD (planes in `prism_ws`, or, when null, in a
// stream-ordered allocation of this call released on `st` after its
// kernels), the register-resident kernel when it applies, else the LDS-tiled
// one.
// epi (optional): the RGB epilogue, fused into the register-resident kernel
// or run as one in-place pass after the demosaic.
int launch_frame(const FrameParams& f, int algo, int mode, void* prism_ws, hipStream_t st,
                 const Epilogue* epi = nullptr);

// A stream-ordered allocation of `bytes` on `st` (the library's pool when
// eager, a graph node under capture); release it with hipFreeAsync on `st`.
hipError_t planes_alloc(size_t bytes, hipStream_t st, void** out);

// Checks, parameters and launch of one frame from device buffers
// (imop_hip_demosaic; with an epilogue, imop_hip_demosaic_isp).
int run_demosaic(const void* d_in, uint64_t in_bytes, void* d_out, uint64_t out_bytes, int w, int h, int pattern,
                 int algo, int bd, int packed, const float* bl, int bl_mode, hipStream_t st,
                 const Epilogue* epi = nullptr);

// Number of visible devices (probed once).
int device_count_cached();

// A batch frame's output element size in bytes (out_elem_size 0 -> 1 for
// depths up to 8, else 2), and imop_hip_demosaic's checks on one batch frame
// plus its output element and buffer sizes.
int out_elem_of(const imop_hip_frame_desc& d);
int check_frame(const imop_hip_frame_desc& d);

}  // namespace imop_hip
