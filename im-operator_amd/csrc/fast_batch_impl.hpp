// fast_batch_impl.hpp -- batched form of the register-resident HQLI / MG
// kernel (fast_impl.hpp).
//
// Work list.  Every frame of the batch is cut into the same workgroup items
// as a single-frame launch: a strip (strip_count; aligned strips for u16
// RGB) by a group row of 4 waves x TH rows.  Items are numbered frame by
// frame (BatchFrame::first), so the list is ragged: frames of any size, bit
// depth, black level and output width share a launch.  Each item runs
// group_item, the single-frame kernel's own code for that group: the
// shared-LDS path (careful columns on edge strips), or per-wave rim bands at
// the frame's top and bottom rows.
//
// One launch per raw layout class (u8 / u16 / 12-bit / 10-bit packed, times
// w % 4 == 0 or 2) present in the batch: a kernel holding one class needs
// 54-64 VGPRs, 8 waves per SIMD, where one inlining all eight classes needs
// 91-123 (4-5 waves per SIMD) and ran the 64-frame mix 1.9x slower.
//
// Schedule.  Workgroups are dealt round-robin to the 8 XCDs.  Frame-major
// placement (order bit 2, the default; one workgroup per item): each frame
// is cut into 8 contiguous runs, one per XCD, exactly as its single-frame
// launch places its groups, and the frames follow one another through the
// grid (frame_major_item).  Without it the whole list is cut into 8
// contiguous slices, one per XCD, so that the XCDs work on frames up to 1.4
// frames apart (the 64-frame mix ran 11 % slower so).  Either way groups
// that share raw rows run on the same XCD (one L2), and XCDs 4-7 walk their
// runs from the end so the last run's final rows do not finish the launch.
// PERSIST = false: one workgroup per item.  PERSIST = true: a grid of a few
// workgroups per CU walks each slice with a stride of that slice's workgroup
// count (its loop keeps 13 more VGPRs live, 6 waves per SIMD).  Each item's
// frame comes from a chunk table (item_frame) by scalar loads, and its strip
// and group row from place_split with the frame's reciprocal, as in the
// single-frame kernel: a batch of one 12 MP frame then runs in the
// single-frame kernel's time.
//
// Wave form (k_batch_w, waves per item 1): every wave is one item, a strip
// by a band of 4 or 8 rows on the single-frame kernel's per-wave path, in
// workgroups of one wave placed like the group form's workgroups.
#pragma once

#include "batch.hpp"
#include "fast_impl.hpp"

namespace imop_hip {

// *slot, a pointer to global memory that is itself read from device memory,
// loaded as a global pointer: addressing through it then compiles to global
// loads and stores, as through the pointers read from the kernel arguments,
// not to flat ones (which wait on both the LDS and the memory counters; the
// 12 MP RAW12 batch item ran 1.2x slower with them).
template <typename T>
__device__ __forceinline__ T* load_global_ptr(T* const* slot) {
    typedef __attribute__((address_space(1))) T gT;
    return (T*)*(gT* const*)(const void*)slot;
}

// The frame of work item `it` (wave-uniform): the last frame with first <=
// it.  One scalar load of its chunk's entry settles it unless a frame starts
// inside the chunk past the chunk's first item; then the frame table is
// scanned forward from the entry's frame.
__device__ __forceinline__ int item_frame(const BatchArgs& a, uint32_t it) {
    const uint2 e = a.chunk[it / kBatchChunk];
    int fi = (int)e.x;
    if (it >= e.y) {
        ++fi;
        while (fi + 1 < a.nframes && a.frames[fi + 1].first <= it) ++fi;
    }
    return fi;
}

// Item `local` (frame-local index) of frame bf, as wave w of the workgroup.
template <int FMT, int PAT, int MODE, int TH, bool ODD>
__device__ __forceinline__ void batch_item_at(const BatchArgs& a, const BatchFrame& bf, uint32_t local, int w,
                                              uint8_t* L) {
    constexpr int GB = 4;
    using G = GroupLayout<TH, GB>;
    FastParams f{};
    f.in = load_global_ptr(&bf.in);
    f.out = load_global_ptr(&bf.out);
    f.in_bytes = bf.in_bytes;
    f.w = bf.w;
    f.h = bf.h;
    f.bd = bf.bd;
    f.maxv = bf.maxv;
    f.clamp_hi = bf.clamp_hi;
    f.pattern = PAT;
    f.out16 = bf.flags & 1;
    f.bl_on = (bf.flags >> 1) & 1;
    for (int k = 0; k < 4; ++k) f.bl_i[k] = bf.bl_i[k];
    const int nsx = bf.nsx, nby = bf.nby;
    f.aligned = 1;  // aligned strips for u16 RGB, as the single-frame kernel
    f.th = TH;
    f.wpb = GB;
    f.nsx = nsx;
    f.nby = nby;
    int sx, gy;
    place_split(local, (a.order & 2) != 0, (uint32_t)nsx, (uint32_t)nby, bf.place_mul, sx, gy);
    group_item<FMT, PAT, MODE, TH, GB, ODD, false, false>(f, strip_of(f, sx), gy * G::kRows, w, L, nullptr);
}

// One work item `it` of a batch launch (item-major list), as wave w.
template <int FMT, int PAT, int MODE, int TH, bool ODD>
__device__ __forceinline__ void batch_item(const BatchArgs& a, uint32_t it, int w, uint8_t* L) {
    const BatchFrame& bf = a.frames[item_frame(a, it)];
    batch_item_at<FMT, PAT, MODE, TH, ODD>(a, bf, it - bf.first, w, L);
}

// Frame-major placement (order bit 2, one workgroup per item): workgroup b
// is slot b >> 3 of XCD b & 7.  Slots are numbered frame by frame
// (BatchFrame::first counts slots, ceil(items / 8) a frame), and in its
// slots of a frame XCD x runs the x-th contiguous eighth of its items
// (frame_major_local): every frame is placed as one single-frame launch
// places its groups, and the frames follow one another through the grid.
// Returns the frame (null for the at most 7 idle workgroups of a frame whose
// item count is not a multiple of 8) and its item in `local`.
__device__ __forceinline__ const BatchFrame* frame_major_item(const BatchArgs& a, uint32_t& local) {
    const uint32_t s = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x >> 3));
    if (s >= a.total) return nullptr;
    const BatchFrame& bf = a.frames[item_frame(a, s)];
    const uint32_t n = (uint32_t)bf.nsx * (uint32_t)bf.nby;
    return frame_major_local(n, blockIdx.x & 7u, s - bf.first, local) ? &bf : nullptr;
}

template <int FMT, int PAT, int MODE, int TH, bool ODD, bool PERSIST>
__global__ __launch_bounds__(256) void k_batch(BatchArgs a) {
    using G = GroupLayout<TH, 4>;
    __shared__ uint4 s_group[G::kBytes / 16];
    uint8_t* L = (uint8_t*)s_group;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nblk = gridDim.x;
    if (!PERSIST && (a.order & 4)) {
        uint32_t local = 0;
        const BatchFrame* bf = frame_major_item(a, local);
        if (bf) batch_item_at<FMT, PAT, MODE, TH, ODD>(a, *bf, local, w, L);
        return;
    }
    if (!PERSIST) {
        // one workgroup per item (the host launches `total` workgroups): the
        // single-frame kernel's bijection, XCD x running the contiguous run
        // of cnt = q + (x < r) items (nblk = 8q + r), XCDs 4-7 from its end
        const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
        const int q8 = nblk >> 3, r8 = nblk & 7;
        const int cnt = q8 + (xcd < r8 ? 1 : 0);
        const int run = xcd >= 4 ? cnt - 1 - slot : slot;
        const uint32_t it = (uint32_t)__builtin_amdgcn_readfirstlane(xcd * q8 + imin(xcd, r8) + run);
        if (it < a.total) batch_item<FMT, PAT, MODE, TH, ODD>(a, it, w, L);
        return;
    }
    // S slices of the list, one per XCD (fewer when the grid has fewer than 8
    // workgroups, so that every slice has at least one workgroup)
    const int S = nblk < 8 ? nblk : 8;
    const int xcd = (int)blockIdx.x % S, slot = (int)blockIdx.x / S;
    const uint32_t nx = (uint32_t)(nblk / S + (xcd < nblk % S ? 1 : 0));  // workgroups on this slice
    const uint32_t lo = (uint32_t)((uint64_t)a.total * (uint32_t)xcd / (uint32_t)S);
    const uint32_t hi = (uint32_t)((uint64_t)a.total * (uint32_t)(xcd + 1) / (uint32_t)S);
    const bool rev = xcd >= 4;
    for (uint32_t j = (uint32_t)slot; j < hi - lo; j += nx) {
        batch_item<FMT, PAT, MODE, TH, ODD>(a, rev ? hi - 1u - j : lo + j, w, L);
        __syncthreads();  // the next item reuses the group's LDS
    }
}

// Wave form: every wave is one work item, a strip by a band of TH rows
// (items numbered frame by frame as nsx x nby bands), on the single-frame
// kernel's per-wave path (run_band); no LDS is shared and no barrier taken.
// Workgroups of one wave, placed like the group form's workgroups.
template <int FMT, int PAT, int MODE, int TH, bool ODD>
__device__ __forceinline__ void batch_wave_item_at(const BatchArgs& a, const BatchFrame& bf, uint32_t local,
                                                   uint32_t* stage) {
    FastParams f{};
    f.in = load_global_ptr(&bf.in);
    f.out = load_global_ptr(&bf.out);
    f.in_bytes = bf.in_bytes;
    f.w = bf.w;
    f.h = bf.h;
    f.bd = bf.bd;
    f.maxv = bf.maxv;
    f.clamp_hi = bf.clamp_hi;
    f.pattern = PAT;
    f.out16 = bf.flags & 1;
    f.bl_on = (bf.flags >> 1) & 1;
    for (int k = 0; k < 4; ++k) f.bl_i[k] = bf.bl_i[k];
    const int nsx = bf.nsx, nby = bf.nby;
    f.aligned = 1;
    f.th = TH;
    f.wpb = 1;
    f.nsx = nsx;
    f.nby = nby;
    int sx, by;
    place_split(local, (a.order & 2) != 0, (uint32_t)nsx, (uint32_t)nby, bf.place_mul, sx, by);
    const Strip sp = strip_of(f, sx);
    const int y0 = by * TH;
    const bool vrim = y0 < 3 || y0 + TH + 3 > f.h - 1;
    const bool hrim = sp.xs < 3 || sp.xs + 248 > f.w - 3;
    if (vrim) run_band<FMT, PAT, MODE, true, TH, ODD, false, true>(f, sp, y0, stage);
    else if (hrim) run_band<FMT, PAT, MODE, true, TH, ODD, false, false>(f, sp, y0, stage);
    else run_band<FMT, PAT, MODE, false, TH, ODD, false, false>(f, sp, y0, stage);
}

template <int FMT, int PAT, int MODE, int TH, bool ODD>
__device__ __forceinline__ void batch_wave_item(const BatchArgs& a, uint32_t it, uint32_t* stage) {
    const BatchFrame& bf = a.frames[item_frame(a, it)];
    batch_wave_item_at<FMT, PAT, MODE, TH, ODD>(a, bf, it - bf.first, stage);
}

template <int FMT, int PAT, int MODE, int TH, bool ODD, bool PERSIST>
__global__ __launch_bounds__(64) void k_batch_w(BatchArgs a) {
    __shared__ uint4 s_stage[128];  // one 1536-byte row slice plus the chunk reads' reach
    uint32_t* stage = (uint32_t*)s_stage;
    const int nblk = gridDim.x;
    if (!PERSIST && (a.order & 4)) {
        uint32_t local = 0;
        const BatchFrame* bf = frame_major_item(a, local);
        if (bf) batch_wave_item_at<FMT, PAT, MODE, TH, ODD>(a, *bf, local, stage);
        return;
    }
    if (!PERSIST) {
        const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
        const int q8 = nblk >> 3, r8 = nblk & 7;
        const int cnt = q8 + (xcd < r8 ? 1 : 0);
        const int run = xcd >= 4 ? cnt - 1 - slot : slot;
        const uint32_t it = (uint32_t)__builtin_amdgcn_readfirstlane(xcd * q8 + imin(xcd, r8) + run);
        if (it < a.total) batch_wave_item<FMT, PAT, MODE, TH, ODD>(a, it, stage);
        return;
    }
    const int S = nblk < 8 ? nblk : 8;
    const int xcd = (int)blockIdx.x % S, slot = (int)blockIdx.x / S;
    const uint32_t nx = (uint32_t)(nblk / S + (xcd < nblk % S ? 1 : 0));
    const uint32_t lo = (uint32_t)((uint64_t)a.total * (uint32_t)xcd / (uint32_t)S);
    const uint32_t hi = (uint32_t)((uint64_t)a.total * (uint32_t)(xcd + 1) / (uint32_t)S);
    const bool rev = xcd >= 4;
    for (uint32_t j = (uint32_t)slot; j < hi - lo; j += nx)
        batch_wave_item<FMT, PAT, MODE, TH, ODD>(a, rev ? hi - 1u - j : lo + j, stage);
}

// k_batch of one layout class over `blocks` workgroups: wpb 4, group items
// of 4 waves x rows (2 or 4); wpb 1, wave items of rows (4 or 8).
template <int FMT, int PAT, int MODE, bool ODD>
hipError_t launch_batch_class(const BatchArgs& a, int rows, bool persist, int blocks, hipStream_t st) {
    const dim3 grid(blocks), block(256);
    if (a.wpb == 1) {
        const dim3 wblock(64);
        if (rows == 8) {
            if (persist) hipLaunchKernelGGL((k_batch_w<FMT, PAT, MODE, 8, ODD, true>), grid, wblock, 0, st, a);
            else hipLaunchKernelGGL((k_batch_w<FMT, PAT, MODE, 8, ODD, false>), grid, wblock, 0, st, a);
        } else {
            if (persist) hipLaunchKernelGGL((k_batch_w<FMT, PAT, MODE, 4, ODD, true>), grid, wblock, 0, st, a);
            else hipLaunchKernelGGL((k_batch_w<FMT, PAT, MODE, 4, ODD, false>), grid, wblock, 0, st, a);
        }
        return hipGetLastError();
    }
    if (rows == 4) {
        if (persist) hipLaunchKernelGGL((k_batch<FMT, PAT, MODE, 4, ODD, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_batch<FMT, PAT, MODE, 4, ODD, false>), grid, block, 0, st, a);
    } else {
        if (persist) hipLaunchKernelGGL((k_batch<FMT, PAT, MODE, 2, ODD, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_batch<FMT, PAT, MODE, 2, ODD, false>), grid, block, 0, st, a);
    }
    return hipGetLastError();
}

template <int FMT, int PAT, int ALG>
hipError_t launch_batch_fpa(const BatchArgs& a, bool odd, int rows, bool persist, int blocks, hipStream_t st) {
    // MG: 8-bit frames take the reference's integer path, deeper ones the exact rational one
    constexpr int MODE = ALG == 0 ? MODE_HQLI : (FMT == FMT_U8 ? MODE_MG_INT8 : MODE_MG_EXACT);
    return odd ? launch_batch_class<FMT, PAT, MODE, true>(a, rows, persist, blocks, st)
               : launch_batch_class<FMT, PAT, MODE, false>(a, rows, persist, blocks, st);
}

// k_batch for Bayer pattern PAT: algorithm alg (0 HQLI, 1 MG), layout fmt.
template <int PAT>
hipError_t launch_batch_p(const BatchArgs& a, int alg, int fmt, bool odd, int rows, bool persist, int blocks,
                          hipStream_t st) {
    switch (fmt) {
    case FMT_U8:
        return alg == 0 ? launch_batch_fpa<FMT_U8, PAT, 0>(a, odd, rows, persist, blocks, st)
                        : launch_batch_fpa<FMT_U8, PAT, 1>(a, odd, rows, persist, blocks, st);
    case FMT_U16:
        return alg == 0 ? launch_batch_fpa<FMT_U16, PAT, 0>(a, odd, rows, persist, blocks, st)
                        : launch_batch_fpa<FMT_U16, PAT, 1>(a, odd, rows, persist, blocks, st);
    case FMT_P12:
        return alg == 0 ? launch_batch_fpa<FMT_P12, PAT, 0>(a, odd, rows, persist, blocks, st)
                        : launch_batch_fpa<FMT_P12, PAT, 1>(a, odd, rows, persist, blocks, st);
    default:
        return alg == 0 ? launch_batch_fpa<FMT_P10, PAT, 0>(a, odd, rows, persist, blocks, st)
                        : launch_batch_fpa<FMT_P10, PAT, 1>(a, odd, rows, persist, blocks, st);
    }
}

}  // namespace imop_hip
