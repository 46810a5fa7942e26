// window_decode.hpp -- raw samples of a rectangular window that lies inside
// the frame, for the LDS-tiled kernels' staging loops (PRISM / RCD pass A,
// AHD, AMAZE, DFPD, L7, SUPER_FAST).  The same values as sample_at
// (device_common.hpp) pixel for pixel, with the per-pixel work it does for
// every layout and position taken out where the window allows: one loop per
// raw layout, 32-bit offsets, no stream-end bound, and each lane's
// black-level offsets chosen once.
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

// Rows of the window for one raw layout: lane c takes columns xs + c,
// xs + c + 64, ... of rows wv, wv + 4, ... (wave wv of 256 threads) and
// hands put(j, c, v) the value sample_at(f, xs + c, ys + j) would return --
// black level, or the u16 MSB alignment without it -- with the lane's two
// black-level offsets (even and odd rows) chosen once (its columns share
// their parity).
template <int FMT, int RW, int RH, class Put>
__device__ __forceinline__ void window_rows(const FrameParams& f, int xs, int ys, Put put) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane >= RW) return;
    const bool xodd = ((xs + lane) & 1) != 0;
    const float bl_even = xodd ? f.bl_q[1] : f.bl_q[0], bl_odd = xodd ? f.bl_q[3] : f.bl_q[2];
    const bool msb = FMT == FMT_U16 && f.bd > 8 && f.bd < 16;
    for (int j = wv; j < RH; j += 4) {
        const int y = ys + j;
        const uint32_t row = (uint32_t)y * (uint32_t)f.w + (uint32_t)xs;
        const float bl = (y & 1) ? bl_odd : bl_even;
        for (int c = lane; c < RW; c += 64) {
            int v = raw_inside<FMT>(f.in, row + (uint32_t)c);
            if (f.bl_on) v = iclamp((int)((float)v - bl), 0, f.maxv);
            else if (msb && v > f.maxv) v >>= (16 - f.bd);
            put(j, c, v);
        }
    }
}

// Windows wider than a wave: element k = t, t + 256, ... of the window in
// row-major order (thread t of 256), so every lane has work on every step;
// the black-level offset chosen per element from its row and column parity.
template <int FMT, int RW, int RH, class Put>
__device__ __forceinline__ void window_elems(const FrameParams& f, int xs, int ys, Put put) {
    const bool msb = FMT == FMT_U16 && f.bd > 8 && f.bd < 16;
    for (int k = (int)threadIdx.x; k < RW * RH; k += 256) {
        const int j = k / RW, c = k - j * RW;
        const int y = ys + j, x = xs + c;
        int v = raw_inside<FMT>(f.in, (uint32_t)y * (uint32_t)f.w + (uint32_t)x);
        if (f.bl_on) {
            const float bl = (y & 1) ? ((x & 1) ? f.bl_q[3] : f.bl_q[2]) : ((x & 1) ? f.bl_q[1] : f.bl_q[0]);
            v = iclamp((int)((float)v - bl), 0, f.maxv);
        } else if (msb && v > f.maxv) {
            v >>= (16 - f.bd);
        }
        put(j, c, v);
    }
}

// Rows when they leave at most a quarter of the lanes of their column steps
// idle (and every window of at most 64 columns), else elements: measured in
// VALU instructions per 12 MP frame, rows took SUPER_FAST (258 columns) to
// 16.7 M and L7 (264) to 23.1 M against 19.2 M and 25.6 M by elements, and
// elements took DFPD (70) to 39.0 M against 42.1 M by rows.
template <int FMT, int RW, int RH, class Put>
__device__ __forceinline__ void window_fill(const FrameParams& f, int xs, int ys, Put put) {
    constexpr int idle = (RW + 63) / 64 * 64 - RW;
    if constexpr (RW <= 64 || 4 * idle <= RW) window_rows<FMT, RW, RH>(f, xs, ys, put);
    else window_elems<FMT, RW, RH>(f, xs, ys, put);
}

// The RW x RH window with top-left frame position (xs, ys) by window_rows
// (at most 64 columns) or window_elems when it lies inside the frame and
// above its last row and the layout is u8 / u16 / 12-bit / 10-bit packed;
// false (nothing done) for the others -- edge windows, the GigE and generic
// bit layouts -- which the caller stages through sample_at.
// Workgroup-uniform; 256 threads.
template <int RW, int RH, class Put>
__device__ __forceinline__ bool window_inside(const FrameParams& f, int xs, int ys, Put put) {
    if (!(xs >= 0 && ys >= 0 && xs + RW <= f.w && ys + RH < f.h)) return false;
    switch (f.fmt) {
    case FMT_U8: window_fill<FMT_U8, RW, RH>(f, xs, ys, put); return true;
    case FMT_U16: window_fill<FMT_U16, RW, RH>(f, xs, ys, put); return true;
    case FMT_P12: window_fill<FMT_P12, RW, RH>(f, xs, ys, put); return true;
    case FMT_P10: window_fill<FMT_P10, RW, RH>(f, xs, ys, put); return true;
    default: return false;
    }
}

}  // namespace imop_hip
