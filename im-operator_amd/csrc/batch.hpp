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
    uint16_t nsx, nby;     // strips (strip_count) and group rows (rows per wave x waves per workgroup)
    uint32_t first;        // index of the frame's first work item in the batch
};
static_assert(sizeof(BatchFrame) == 72, "BatchFrame layout");

// Items per entry of the chunk table: chunk c's entry is the frame holding
// item c * kBatchChunk, so an item's frame is that entry or a later one.
constexpr int kBatchChunk = 32;

// Kernel arguments of one batched launch.
struct BatchArgs {
    const BatchFrame* frames;  // device table, sorted by `first`
    const uint16_t* chunk_frame;  // device table: frame of item c * kBatchChunk, ceil(total / kBatchChunk) entries
    int32_t nframes;
    uint32_t total;            // work items (frame, strip, group row) of all frames
    int32_t wpb;               // waves per workgroup (stacked down the strip)
    int32_t order;             // bit 1: strip-fastest within a frame
};

// Enqueue k_batch for one layout class -- pattern, alg (0 HQLI, 1 MG), raw
// layout fmt (FMT_U8 / U16 / P12 / P10), odd (w % 4 == 2) -- with `rows`
// rows per wave (2 or 4), over `blocks` workgroups of 4 waves: one per item,
// or a persistent grid walking the list.
hipError_t launch_batch(const BatchArgs& a, int pattern, int alg, int fmt, bool odd, int rows, bool persist,
                        int blocks, hipStream_t st);

}  // namespace imop_hip
