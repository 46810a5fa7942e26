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
    uint32_t first;        // the frame's first work item in the batch (frame-major placement: its first slot)
    uint64_t place_mul;    // place_reciprocal(nsx * nby, strip-fastest ? nsx : nby)
};
static_assert(sizeof(BatchFrame) == 80, "BatchFrame layout");

// Items per entry of the chunk table.  Chunk c's entry {fi, end}: fi is the
// frame holding item c * kBatchChunk and end the first item of frame fi + 1
// (the batch's item count for the last frame), so an item of the chunk below
// end belongs to fi and one at or past it to a later frame.
constexpr int kBatchChunk = 32;

// Frame-major placement of one frame of n items: slot j (counted from the
// frame's first slot, j < ceil(n / 8)) of XCD xcd runs item `local`, where
// XCD x owns the contiguous run of q + (x < r) items (n = 8q + r) after
// those of XCDs 0..x-1 and XCDs 4-7 walk their runs from the end -- the
// single-frame kernel's placement of a launch of n workgroups.  False for
// the idle slot of an XCD whose run is one short.
__host__ __device__ inline bool frame_major_local(uint32_t n, uint32_t xcd, uint32_t j, uint32_t& local) {
    const uint32_t q = n >> 3, r = n & 7u, cnt = q + (xcd < r ? 1u : 0u);
    if (j >= cnt) return false;
    local = xcd * q + (xcd < r ? xcd : r) + (xcd >= 4 ? cnt - 1u - j : j);
    return true;
}

// Kernel arguments of one batched launch.
struct BatchArgs {
    const BatchFrame* frames;  // device table, sorted by `first`
    const uint2* chunk;        // device table of ceil(total / kBatchChunk) entries {fi, end}
    int32_t nframes;
    uint32_t total;            // work items (frame, strip, group row) of all frames; frame-major: slots
    int32_t wpb;               // 4: group items (4 waves down a strip); 1: one wave per item
    int32_t order;             // bit 1: strip-fastest within a frame; bit 2: frame-major placement
};

// Enqueue k_batch for one layout class -- pattern, alg (0 HQLI, 1 MG), raw
// layout fmt (FMT_U8 / U16 / P12 / P10), odd (w % 4 == 2) -- with `rows`
// rows per wave (group items of a.wpb = 4 waves: 2 or 4; wave items, a.wpb =
// 1: 4 or 8), over `blocks` workgroups: one per item, or a persistent grid
// walking the list.
hipError_t launch_batch(const BatchArgs& a, int pattern, int alg, int fmt, bool odd, int rows, bool persist,
                        int blocks, hipStream_t st);

}  // namespace imop_hip
