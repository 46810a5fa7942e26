// batch.hpp -- the batched form of the register-resident HQLI / MG kernel:
// many frames of any size, layout and bit depth in one persistent launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace imop_hip {

// One frame of a batch, as the kernel reads it from the device-side table
// (scalar loads; 64 bytes).
struct BatchFrame {
    const uint8_t* in;
    uint8_t* out;
    uint32_t in_bytes;     // packed stream / unpacked buffer bytes (< 2^31)
    int32_t w, h;
    int32_t bd, maxv, clamp_hi;
    int32_t fmt;           // FMT_U8 / FMT_U16 / FMT_P12 / FMT_P10
    int32_t flags;         // bit 0: RGB as u16; bit 1: black level applied
    int32_t bl_i[4];       // integral black level by (y & 1) * 2 + (x & 1); 0 without
    uint16_t nsx, nby;     // strips (strip_count) and bands at the batch's rows per wave
    uint32_t first;        // index of the frame's first work item in the batch
};
static_assert(sizeof(BatchFrame) == 72, "BatchFrame layout");

// Kernel arguments of one batched launch.
struct BatchArgs {
    const BatchFrame* frames;  // device table, sorted by `first`
    int32_t nframes;
    uint32_t total;            // work items (frame, strip, band) of all frames
    int32_t wpb;               // waves per workgroup
    int32_t order;             // bit 1: strip-fastest within a frame
};

// Enqueue k_batch<pattern, alg (0 HQLI, 1 MG), rows> over `blocks` workgroups.
hipError_t launch_batch(const BatchArgs& a, int pattern, int alg, int rows, int blocks, hipStream_t st);

}  // namespace imop_hip
