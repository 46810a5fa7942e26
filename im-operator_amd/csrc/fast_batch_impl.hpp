// fast_batch_impl.hpp -- persistent batched form of the register-resident
// HQLI / MG kernel (fast_impl.hpp).
//
// Work list.  Every frame of the batch is cut into the same work items as a
// single-frame launch: one wave-sized strip (strip_count) by a band of TH
// rows.  Items are numbered frame by frame (BatchFrame::first), so the list
// is ragged: frames of any size, layout (u8, u16, 12-bit, 10-bit packed),
// width class (w % 4 == 0 or 2), bit depth, black level and output width
// share one launch.
//
// Schedule.  Workgroups are dealt round-robin to the 8 XCDs; the list is cut
// into 8 contiguous slices and the waves of XCD x walk slice x with a stride
// of that slice's wave count, so bands that share a vertical apron run on the
// same XCD (one L2) and the grid can be sized independently of the work
// (persistent: a few workgroups per CU loop until their slice is done).
// Each wave finds the frame of its first item by binary search over the
// sorted table and of later items by scanning forward (scalar loads).
#pragma once

#include "batch.hpp"
#include "fast_impl.hpp"

namespace imop_hip {

namespace {

template <int FMT, int PAT, int MODE, int TH, bool ODD>
__device__ __forceinline__ void batch_item(const FastParams& f, const Strip& sp, int y0, bool careful,
                                           uint32_t* stage) {
    if (careful) run_band<FMT, PAT, MODE, true, TH, ODD>(f, sp, y0, stage);
    else run_band<FMT, PAT, MODE, false, TH, ODD>(f, sp, y0, stage);
}

template <int FMT, int PAT, int ALG, int TH>
__device__ __forceinline__ void batch_fmt(const FastParams& f, const Strip& sp, int y0, bool careful,
                                          uint32_t* stage) {
    // MG: 8-bit frames take the reference's integer path, deeper ones the exact rational one
    constexpr int MODE = ALG == 0 ? MODE_HQLI : (FMT == FMT_U8 ? MODE_MG_INT8 : MODE_MG_EXACT);
    if (f.w & 2) batch_item<FMT, PAT, MODE, TH, true>(f, sp, y0, careful, stage);
    else batch_item<FMT, PAT, MODE, TH, false>(f, sp, y0, careful, stage);
}

}  // namespace

template <int PAT, int ALG, int TH>
__global__ __launch_bounds__(256) void k_batch(BatchArgs a) {
    __shared__ uint4 s_stage[4][96];  // one 1536-byte row slice per wave
    uint32_t* stage = (uint32_t*)s_stage[threadIdx.x >> 6];
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // S slices, one per XCD (fewer when the grid has fewer than 8
    // workgroups, so that every slice has at least one workgroup)
    const int nblk = gridDim.x;
    const int S = nblk < 8 ? nblk : 8;
    const int xcd = (int)blockIdx.x % S, slot = (int)blockIdx.x / S;
    const int nx = nblk / S + (xcd < nblk % S ? 1 : 0);  // workgroups on this slice
    const uint32_t lo = (uint32_t)((uint64_t)a.total * (uint32_t)xcd / (uint32_t)S);
    const uint32_t hi = (uint32_t)((uint64_t)a.total * (uint32_t)(xcd + 1) / (uint32_t)S);
    const uint32_t stride = (uint32_t)(nx * a.wpb);
    uint32_t it = lo + (uint32_t)(slot * a.wpb + wib);
    // frame of the first item: binary search (last frame with first <= it)
    int fi = 0;
    if (it < hi) {
        int f_hi = a.nframes - 1;
        while (fi < f_hi) {
            const int mid = (fi + f_hi + 1) >> 1;
            if (a.frames[mid].first <= it) fi = mid;
            else f_hi = mid - 1;
        }
    }
    for (; it < hi; it += stride) {
        while (fi + 1 < a.nframes && a.frames[fi + 1].first <= it) ++fi;
        const BatchFrame& bf = a.frames[fi];
        FastParams f;
        f.in = bf.in;
        f.out = bf.out;
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
        f.aligned = 0;  // plain strips (strip_stride)
        f.th = TH;
        f.nsx = nsx;
        f.nby = nby;
        const int local = (int)(it - bf.first);
        const bool strip_fastest = (a.order & 2) != 0;
        const int by = strip_fastest ? local / nsx : local % nby;
        const int sx = strip_fastest ? local % nsx : local / nby;
        const Strip sp = strip_of(f, sx);
        const int y0 = by * TH;
        const bool careful = y0 < 3 || y0 + TH + 3 > f.h - 1 || sp.xs < 3 || sp.xs + 248 > f.w - 3;
        switch (bf.fmt) {
        case FMT_U8: batch_fmt<FMT_U8, PAT, ALG, TH>(f, sp, y0, careful, stage); break;
        case FMT_U16: batch_fmt<FMT_U16, PAT, ALG, TH>(f, sp, y0, careful, stage); break;
        case FMT_P12: batch_fmt<FMT_P12, PAT, ALG, TH>(f, sp, y0, careful, stage); break;
        default: batch_fmt<FMT_P10, PAT, ALG, TH>(f, sp, y0, careful, stage); break;
        }
    }
}

template <int PAT>
hipError_t launch_batch_p(const BatchArgs& a, int alg, int rows, int blocks, hipStream_t st) {
    const dim3 grid(blocks), block(64 * a.wpb);
    if (alg == 0) {
        if (rows == 8) hipLaunchKernelGGL((k_batch<PAT, 0, 8>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_batch<PAT, 0, 4>), grid, block, 0, st, a);
    } else {
        if (rows == 8) hipLaunchKernelGGL((k_batch<PAT, 1, 8>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_batch<PAT, 1, 4>), grid, block, 0, st, a);
    }
    return hipGetLastError();
}

}  // namespace imop_hip
