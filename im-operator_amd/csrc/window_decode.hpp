// window_decode.hpp -- raw samples of a rectangular window that lies inside
// the frame, for the LDS-tiled kernels' staging loops (PRISM / RCD pass A,
// AHD).  The same values as sample_at (device_common.hpp) pixel for pixel,
// with the per-pixel work it does for every layout and position taken out
// where the window allows: one loop per raw layout, 32-bit offsets, no
// stream-end bound, and each lane's black-level offsets chosen once.
#pragma once

#include "device_common.hpp"

namespace imop_hip {

// Raw sample at pixel index i of a window that lies inside the frame and
// above its last row: the decoders of sample_at without the stream-end
// bound, which only the last row's final group can reach, in 32-bit offsets
// (callers take frames of fewer than 2^31 pixels).
template <int FMT>
__device__ __forceinline__ int raw_inside(const uint8_t* __restrict__ d, uint32_t i) {
    if (FMT == FMT_U8) return d[i];
    if (FMT == FMT_U16) return ((const uint16_t*)d)[i];
    if (FMT == FMT_P12) {
        const uint32_t o = (i >> 1) * 3u;
        const unsigned b1 = d[o + 1];
        return (i & 1u) == 0 ? (int)(((unsigned)d[o] << 4) | (b1 & 0xFu)) : (int)(((unsigned)d[o + 2] << 4) | (b1 >> 4));
    }
    const uint32_t o = (i >> 2) * 5u, sub = i & 3u;  // FMT_P10
    return (int)(((unsigned)d[o + sub] << 2) | (((unsigned)d[o + 4] >> (2u * sub)) & 3u));
}

// Rows of the window for one raw layout: lane c takes column xs + c of rows
// wv, wv + 4, ... (wave wv of 256 threads) and hands put(j, c, v) the value
// sample_at(f, xs + c, ys + j) would return -- black level, or the u16 MSB
// alignment without it -- with the lane's two black-level offsets (even and
// odd rows) chosen once.
template <int FMT, int RW, int RH, class Put>
__device__ __forceinline__ void window_rows(const FrameParams& f, int xs, int ys, Put put) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane >= RW) return;
    const int x = xs + lane;
    const bool xodd = (x & 1) != 0;
    const float bl_even = xodd ? f.bl_q[1] : f.bl_q[0], bl_odd = xodd ? f.bl_q[3] : f.bl_q[2];
    const bool msb = FMT == FMT_U16 && f.bd > 8 && f.bd < 16;
    for (int j = wv; j < RH; j += 4) {
        const int y = ys + j;
        int v = raw_inside<FMT>(f.in, (uint32_t)y * (uint32_t)f.w + (uint32_t)x);
        if (f.bl_on) v = iclamp((int)((float)v - ((y & 1) ? bl_odd : bl_even)), 0, f.maxv);
        else if (msb && v > f.maxv) v >>= (16 - f.bd);
        put(j, lane, v);
    }
}

// The RW x RH window with top-left frame position (xs, ys), RW <= 64, by
// window_rows when it lies inside the frame and above its last row and the
// layout is u8 / u16 / 12-bit / 10-bit packed; false (nothing done) for the
// others -- edge windows, the GigE and generic bit layouts -- which the
// caller stages through sample_at.  Workgroup-uniform.
template <int RW, int RH, class Put>
__device__ __forceinline__ bool window_inside(const FrameParams& f, int xs, int ys, Put put) {
    static_assert(RW <= 64, "window_inside: one row per wave step");
    if (!(xs >= 0 && ys >= 0 && xs + RW <= f.w && ys + RH < f.h)) return false;
    switch (f.fmt) {
    case FMT_U8: window_rows<FMT_U8, RW, RH>(f, xs, ys, put); return true;
    case FMT_U16: window_rows<FMT_U16, RW, RH>(f, xs, ys, put); return true;
    case FMT_P12: window_rows<FMT_P12, RW, RH>(f, xs, ys, put); return true;
    case FMT_P10: window_rows<FMT_P10, RW, RH>(f, xs, ys, put); return true;
    default: return false;
    }
}

}  // namespace imop_hip
