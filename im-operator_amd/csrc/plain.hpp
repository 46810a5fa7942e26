// plain.hpp -- launchers shared between the kernel files and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include "device_common.hpp"

namespace imop_hip {

// HQLI / MG tiled kernel for one frame (plain_kernels.hip).
hipError_t launch_plain(const FrameParams& f, int mode, hipStream_t st);

// Register-resident HQLI / MG kernel (fast_kernels.hip): usable when
// fast_path_ok(); launch shape from set_launch_shape().
bool fast_path_ok(const FrameParams& f);
// epi (optional): RGB stages applied in the same launch (bands of 4 rows).
hipError_t launch_fast(const FrameParams& f, int mode, hipStream_t st, const Epilogue* epi = nullptr,
                       const LensParams* lens = nullptr);
void set_launch_shape(int rows, int wpb, int xcd_remap);
void get_launch_shape(int* rows, int* wpb, int* xcd_remap);

// In-place black level on an unpacked frame; q = offsets by (y & 1) * 2 + (x & 1).
hipError_t launch_black_level(uint8_t* data, int w, int h, int bd, const float q[4], hipStream_t st);

// SUPER_FAST for one frame (super_fast_kernels.hip).
hipError_t launch_super_fast(const FrameParams& f, hipStream_t st);

// L7 for one frame, interior then the width-3 border fill (l7_kernels.hip).
hipError_t launch_l7(const FrameParams& f, hipStream_t st);
// fill_rgb_borders of width b on an output whose interior [m, w-m) x [m, h-m)
// is written, m <= b (l7_kernels.hip).
hipError_t launch_rgb_borders(const FrameParams& f, int b, int m, hipStream_t st);
// DFPD for one frame, interior then the width-5 border fill (dfpd_kernels.hip).
hipError_t launch_dfpd(const FrameParams& f, hipStream_t st);

// AHD for one frame (ahd_kernels.hip).
hipError_t launch_ahd(const FrameParams& f, hipStream_t st);

// AMAZE for one frame (amaze_kernels.hip).
hipError_t launch_amaze(const FrameParams& f, hipStream_t st);

// PRISM for one frame (prism_kernels.hip).  `ws` holds prism_workspace_bytes().
size_t prism_workspace_bytes(int w, int h);
hipError_t launch_prism(const FrameParams& f, void* ws, hipStream_t st, bool rcd = false);

}  // namespace imop_hip
