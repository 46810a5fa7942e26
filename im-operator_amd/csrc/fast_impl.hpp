// fast_impl.hpp -- register-resident fused unpack -> black level ->
// HQLI / MG demosaic for gfx950.
//
// Geometry.  A wave owns a strip of 64 lanes x 4 pixels = 256 columns and a
// band of TH rows.  Lanes 1..62 compute 248 columns; lanes 0 and 63 only
// load, so that every computing lane finds its horizontal neighbours in the
// adjacent lanes (DPP wave_shr / wave_shl, one VALU op per value).  Strips
// advance 248 columns, or, for u16 output of the single-frame kernel, 224
// with each row's strip boundaries shifted onto 64-byte boundaries of the
// output (strip_stride, row_shift), each wave writing a window of the columns
// it computed.
//
// Each lane walks down its band two rows (one CFA quad row) per step, keeping
// a sliding window of 6 raw rows and 4 rows of MG colour differences in
// registers, so every packed byte is loaded once per band (plus a 6-row
// apron) with one aligned 4- or 8-byte load per lane and row, and every
// output row is written once: in interior bands regrouped through a 1.5 KB
// LDS slice per wave into lane-contiguous 16-byte stores, in edge bands with
// direct per-lane stores.  No barriers across waves.
//
// Arithmetic is the exact integer form of the reference algorithms (see
// plain_kernels.hip): MG in units of 1/32 (float path) or the reference's own
// 8-bit integer rounding; HQLI integer averages.  The border rules of
// fill_intermediate_borders / fill_rgb_borders (common.hpp:47-129) are applied
// in the waves that touch the frame edge ("careful" bands): the pre-fill value
// of every position is computed, then rows 0-1 / h-2..h-1 copy rows 2 / h-3
// as they are before the column fill, and columns 0-1 / w-2..w-1 copy
// columns 2 / w-3.
//
// Preconditions (checked on the host, else the generic LDS path runs):
// width % 4 == 0 or 2, layout u8 / u16 / 12-bit packed / 10-bit packed,
// black level absent or integral per site.
#pragma once

#include <type_traits>

#include "device_common.hpp"
#include "epilogue.hpp"
#include "plain.hpp"

namespace imop_hip {

namespace {

// (lanes without a neighbour read 0; only halo lanes see that)
__device__ __forceinline__ int from_left(int v) {   // lane i <- lane i-1
    return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ int from_right(int v) {  // lane i <- lane i+1
    return __builtin_amdgcn_mov_dpp(v, 0x130, 0xF, 0xF, true);
}

// clamp(v, 0, hi) in one v_med3_i32.  The compiler forms it by itself only
// for a constant hi (it cannot prove 0 <= hi for a runtime bound, which is a
// wave-uniform clamp_hi here, hence the scalar operand).
__device__ __forceinline__ int clamp0(int v, int hi) {
    if (__builtin_constant_p(hi)) return imin(imax(v, 0), hi);
    int r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(v), "s"(hi));
    return r;
}

__device__ __forceinline__ int tdiv4(int v) { return (v + ((v >> 31) & 3)) >> 2; }           // C's v / 4
__device__ __forceinline__ int tdiv2(int v) { return (v + (int)((unsigned)v >> 31)) >> 1; }  // C's v / 2

// Raw bytes of one row chunk of a lane: 4 samples at x (x % 4 == 0).  All
// loads of a band are issued back to back with no control flow between
// them: rows outside the frame read offset 0 and are zeroed when decoded.
struct Raw {
    uint32_t lo, hi;
};

// Byte offset of the group holding sample (x, y) = row part + lane part
// (x % 4 == 0 and w % 4 == 0 keep both parts whole numbers of bytes).
template <int FMT>
__device__ __forceinline__ uint32_t row_stride(uint32_t w) {
    return FMT == FMT_U8 ? w : FMT == FMT_U16 ? 2u * w : (FMT == FMT_P12 || FMT == FMT_G10) ? w / 2u * 3u : w / 4u * 5u;
}
template <int FMT>
__device__ __forceinline__ uint32_t lane_offset(uint32_t x) {
    return FMT == FMT_U8 ? x : FMT == FMT_U16 ? 2u * x : (FMT == FMT_P12 || FMT == FMT_G10) ? x / 2u * 3u : x / 4u * 5u;
}

// Edge bands: exact-length loads that never read past the group run, so
// the stream's final row needs no special case (with w % 4 == 0 every
// packed row ends on a whole group, and 12-bit runs start on an even byte).
template <int FMT>
__device__ __forceinline__ Raw load_raw_exact(const FastParams& f, uint32_t off, bool ok) {
    Raw r;
    const uint8_t* p = f.in + (ok ? off : 0u);
    if (FMT == FMT_P12 || FMT == FMT_G10) {
        const uint32_t a = *(const uint16_t*)p, b = *(const uint16_t*)(p + 2), c = *(const uint16_t*)(p + 4);
        r.lo = a | (b << 16);
        r.hi = c;
    } else {
        r.lo = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
        r.hi = p[4];
    }
    return r;
}

template <int FMT>
__device__ __forceinline__ Raw load_raw(const FastParams& f, uint32_t off, bool ok) {
    Raw r;
    if (FMT == FMT_U8) {
        r.lo = *(const uint32_t*)(f.in + (ok ? off : 0u));
        r.hi = 0;
    } else if (FMT == FMT_U16) {
        uint2 v = *(const uint2*)(f.in + (ok ? off : 0u));
        r.lo = v.x; r.hi = v.y;
    } else {
        uint2 v = *(const uint2*)(f.in + (ok ? (off & ~3u) : 0u));
        const uint32_t s = off & 3u;
        r.lo = __builtin_amdgcn_alignbyte(v.y, v.x, s);  // bytes 0..3 of the group run
        r.hi = v.y >> (8u * s);                           // bytes 4..
    }
    return r;
}

// Frames with w % 4 == 2: rows start at any byte, so the group run of a
// lane starts at the general byte offset of pixel i = y * w + x (x % 4 == 0,
// so i % 4 is 0 on even rows and 2 on odd rows).  A 10-bit run that starts at
// sub-pixel 2 begins 2 bytes into its group: [P2hi P3hi tails | P0hi P1hi .. tails].
template <int FMT>
__device__ __forceinline__ uint32_t run_offset(uint32_t w, uint32_t x, uint32_t y) {
    const uint32_t i = y * w + x;
    return FMT == FMT_U8 ? i : FMT == FMT_U16 ? 2u * i : (FMT == FMT_P12 || FMT == FMT_G10) ? (i >> 1) * 3u
                                                                            : (i >> 2) * 5u + (i & 2u);
}

// Bytes off .. off+7 (u8: off .. off+3) from whole aligned words: reads at
// most 11 bytes past off, so callers use it only where the stream continues
// that far (every row but the last).
template <int FMT>
__device__ __forceinline__ Raw load_raw_any(const FastParams& f, uint32_t off) {
    Raw r;
    const uint32_t s = off & 3u;
    const uint32_t* p = (const uint32_t*)(f.in + (off & ~3u));
    if (FMT == FMT_U16) {  // off % 4 == 0
        r.lo = p[0];
        r.hi = p[1];
    } else if (FMT == FMT_U8) {
        const uint32_t d0 = p[0], d1 = p[1];
        r.lo = __builtin_amdgcn_alignbyte(d1, d0, s);
        r.hi = 0;
    } else {
        const uint32_t d0 = p[0], d1 = p[1], d2 = p[2];
        r.lo = __builtin_amdgcn_alignbyte(d1, d0, s);
        r.hi = __builtin_amdgcn_alignbyte(d2, d1, s);
    }
    return r;
}

// Raw value of pixel (x, y), x < w, through the bounded per-pixel decoders
// (a group reaching past the stream decodes as 0): the last row of frames
// with w % 4 == 2, whose final group may be cut short.
template <int FMT>
__device__ __forceinline__ uint32_t raw_pixel(const FastParams& f, uint32_t x, uint32_t y) {
    const uint64_t i = (uint64_t)y * (uint32_t)f.w + x;
    if (FMT == FMT_U8) return f.in[i];
    if (FMT == FMT_U16) return ((const uint16_t*)f.in)[i];
    if (FMT == FMT_P12) return (uint32_t)decode_p12(f.in, i, f.in_bytes);
    if (FMT == FMT_G10) return (uint32_t)decode_g10(f.in, i, f.in_bytes);
    return (uint32_t)decode_p10(f.in, i, f.in_bytes);
}

// Decode a raw chunk to 4 raw values.  sub2: a 10-bit run starting at
// sub-pixel 2 (odd rows of w % 4 == 2 frames).
template <int FMT>
__device__ __forceinline__ void raw_values(Raw r, bool sub2, uint32_t v[4]) {
    if (FMT == FMT_P10 && sub2) {
        v[0] = ((r.lo & 0xFF) << 2) | ((r.lo >> 20) & 3);
        v[1] = (((r.lo >> 8) & 0xFF) << 2) | ((r.lo >> 22) & 3);
        v[2] = ((r.lo >> 24) << 2) | ((r.hi >> 24) & 3);
        v[3] = ((r.hi & 0xFF) << 2) | ((r.hi >> 26) & 3);
        return;
    }
    if (FMT == FMT_U8) {
        v[0] = r.lo & 0xFF; v[1] = (r.lo >> 8) & 0xFF; v[2] = (r.lo >> 16) & 0xFF; v[3] = r.lo >> 24;
    } else if (FMT == FMT_U16) {
        v[0] = r.lo & 0xFFFF; v[1] = r.lo >> 16; v[2] = r.hi & 0xFFFF; v[3] = r.hi >> 16;
    } else if (FMT == FMT_P12) {
        const uint32_t b1 = (r.lo >> 8) & 0xFF, b4 = r.hi & 0xFF;
        v[0] = ((r.lo & 0xFF) << 4) | (b1 & 0xF);
        v[1] = (((r.lo >> 16) & 0xFF) << 4) | (b1 >> 4);
        v[2] = ((r.lo >> 24) << 4) | (b4 & 0xF);
        v[3] = (((r.hi >> 8) & 0xFF) << 4) | (b4 >> 4);
    } else if (FMT == FMT_G10) {
        // GigE 10-bit pairs: P0 = b0 << 2 | b1 & 3, P1 = b2 << 2 | (b1 >> 4) & 3
        const uint32_t b1 = (r.lo >> 8) & 0xFF, b4 = r.hi & 0xFF;
        v[0] = ((r.lo & 0xFF) << 2) | (b1 & 3);
        v[1] = (((r.lo >> 16) & 0xFF) << 2) | ((b1 >> 4) & 3);
        v[2] = ((r.lo >> 24) << 2) | (b4 & 3);
        v[3] = (((r.hi >> 8) & 0xFF) << 2) | ((b4 >> 4) & 3);
    } else {
        const uint32_t t = r.hi & 0xFF;
        v[0] = ((r.lo & 0xFF) << 2) | (t & 3);
        v[1] = (((r.lo >> 8) & 0xFF) << 2) | ((t >> 2) & 3);
        v[2] = (((r.lo >> 16) & 0xFF) << 2) | ((t >> 4) & 3);
        v[3] = ((r.lo >> 24) << 2) | (t >> 6);
    }
}

// 12-bit packed run (bytes b0..b5 = samples P0..P3, P0 = b0 << 4 | b1 & 15,
// P1 = b2 << 4 | b1 >> 4, P2 = b3 << 4 | b4 & 15, P3 = b5 << 4 | b4 >> 4) to
// black-levelled samples, two at a time in 16-bit halves: byte permutes
// gather (b0, b3), (b1, b4) and (b1 b2, b4 b5); X = (P0, P2), Y = (P1, P3);
// the saturating 16-bit subtract is max(0, v - offset) for v < 4096.
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void p12_samples(Raw r, int o0, int o1, uint32_t v[4]) {
    const uint32_t a = __builtin_amdgcn_perm(r.hi, r.lo, 0x0C030C00u);  // b0, 0, b3, 0
    const uint32_t b = __builtin_amdgcn_perm(r.hi, r.lo, 0x0C040C01u);  // b1, 0, b4, 0
    const uint32_t c = __builtin_amdgcn_perm(r.hi, r.lo, 0x05040201u);  // b1, b2, b4, b5
    const u16x2 x = __builtin_elementwise_sub_sat(__builtin_bit_cast(u16x2, (a << 4) | (b & 0x000F000Fu)),
                                                  (u16x2){(uint16_t)o0, (uint16_t)o0});
    const u16x2 y = __builtin_elementwise_sub_sat(__builtin_bit_cast(u16x2, c) >> (u16x2){4, 4},
                                                  (u16x2){(uint16_t)o1, (uint16_t)o1});
    v[0] = x.x; v[1] = y.x; v[2] = x.y; v[3] = y.y;
}

// GigE 10-bit run (b0..b5 = P0..P3, P0 = b0 << 2 | b1 & 3,
// P1 = b2 << 2 | (b1 >> 4) & 3, P2 = b3 << 2 | b4 & 3, P3 = b5 << 2 | (b4 >> 4) & 3)
// the same way: X = (P0, P2) from (b0, b3) and (b1, b4), Y = (P1, P3) from
// (b2, b5) and (b1, b4) >> 4; samples < 1024, so the halves never carry.
__device__ __forceinline__ void g10_samples(Raw r, int o0, int o1, uint32_t v[4]) {
    const uint32_t a = __builtin_amdgcn_perm(r.hi, r.lo, 0x0C030C00u);  // b0, 0, b3, 0
    const uint32_t b = __builtin_amdgcn_perm(r.hi, r.lo, 0x0C040C01u);  // b1, 0, b4, 0
    const uint32_t c = __builtin_amdgcn_perm(r.hi, r.lo, 0x0C050C02u);  // b2, 0, b5, 0
    const u16x2 x = __builtin_elementwise_sub_sat(__builtin_bit_cast(u16x2, (a << 2) | (b & 0x00030003u)),
                                                  (u16x2){(uint16_t)o0, (uint16_t)o0});
    const u16x2 y = __builtin_elementwise_sub_sat(__builtin_bit_cast(u16x2, (c << 2) | ((b >> 4) & 0x00030003u)),
                                                  (u16x2){(uint16_t)o1, (uint16_t)o1});
    v[0] = x.x; v[1] = y.x; v[2] = x.y; v[3] = y.y;
}

// Black level (integral offsets o0 / o1 for even / odd columns of the row)
// or MSB alignment applied to 4 raw values.
template <int FMT, bool MASK, bool LENS = false>
__device__ __forceinline__ void finish4(const FastParams& f, uint32_t v[4], bool ok, int o0, int o1, int p[4]) {
    if (FMT != FMT_U16) {
        // decoded samples never exceed max; offsets are 0 without black level
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = __builtin_elementwise_sub_sat(v[k], (uint32_t)((k & 1) ? o1 : o0));
    } else if (f.bl_on) {
        // raw u16 may exceed max: black level clamps both ways
#pragma unroll
        for (int k = 0; k < 4; ++k)
            v[k] = (uint32_t)imin((int)__builtin_elementwise_sub_sat(v[k], (uint32_t)((k & 1) ? o1 : o0)), f.maxv);
    } else if (f.bd < 16 && !LENS) {
        // (with lens shading the sample goes to it as stored: the lens stage
        // reads the unpacked buffer and clamps, polynomial.cpp:44-47)
        const int sh = 16 - f.bd;
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = (int)v[k] > f.maxv ? v[k] >> sh : v[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) p[k] = (!MASK || ok) ? (int)v[k] : 0;
}

// Colour role of CFA position parity (py, px) for the pattern.
template <int PAT>
__device__ constexpr int role(int py, int px) { return site_of(PAT, py, px); }

// Column parity of the R/B sites in a row of parity py.
template <int PAT>
__device__ constexpr int rb_col_parity(int py) { return py == r_row_of(PAT) ? r_col_of(PAT) : 1 - r_col_of(PAT); }

// Store the RGB pixels of one lane run at (x, y) (x % 4 == 0); 32-bit
// offsets.  ODD (w % 4 == 2): u16 runs are 4-byte aligned and u8 runs on odd
// rows only 2-byte aligned, and the last lane of a row holds 2 pixels.
template <bool ODD>
__device__ __forceinline__ void store4(const FastParams& f, uint32_t x, uint32_t y, const int r[4], const int g[4],
                                       const int b[4]) {
    const uint32_t i = (y * (uint32_t)f.w + x) * 3u;
    const bool full = !ODD || x + 4u <= (uint32_t)f.w;
    if (f.out16) {
        const uint32_t w0 = (uint32_t)r[0] | ((uint32_t)g[0] << 16), w1 = (uint32_t)b[0] | ((uint32_t)r[1] << 16);
        const uint32_t w2 = (uint32_t)g[1] | ((uint32_t)b[1] << 16), w3 = (uint32_t)r[2] | ((uint32_t)g[2] << 16);
        const uint32_t w4 = (uint32_t)b[2] | ((uint32_t)r[3] << 16), w5 = (uint32_t)g[3] | ((uint32_t)b[3] << 16);
        if (!ODD) {
            uint2* o = (uint2*)(f.out + 2u * i);
            o[0] = make_uint2(w0, w1);
            o[1] = make_uint2(w2, w3);
            o[2] = make_uint2(w4, w5);
        } else {
            uint32_t* o = (uint32_t*)(f.out + 2u * i);
            o[0] = w0; o[1] = w1; o[2] = w2;
            if (full) { o[3] = w3; o[4] = w4; o[5] = w5; }
        }
    } else {
        const uint32_t w0 = (uint32_t)r[0] | ((uint32_t)g[0] << 8) | ((uint32_t)b[0] << 16) | ((uint32_t)r[1] << 24);
        const uint32_t w1 = (uint32_t)g[1] | ((uint32_t)b[1] << 8) | ((uint32_t)r[2] << 16) | ((uint32_t)g[2] << 24);
        const uint32_t w2 = (uint32_t)b[2] | ((uint32_t)r[3] << 8) | ((uint32_t)g[3] << 16) | ((uint32_t)b[3] << 24);
        if (!ODD) {
            uint32_t* o = (uint32_t*)(f.out + i);
            o[0] = w0; o[1] = w1; o[2] = w2;
        } else if ((y & 1u) == 0) {  // 4-byte aligned
            uint32_t* o = (uint32_t*)(f.out + i);
            o[0] = w0;
            if (full) { o[1] = w1; o[2] = w2; }
            else *(uint16_t*)(f.out + i + 4) = (uint16_t)w1;
        } else {                     // 2-byte aligned
            uint16_t* o = (uint16_t*)(f.out + i);
            o[0] = (uint16_t)w0; o[1] = (uint16_t)(w0 >> 16); o[2] = (uint16_t)w1;
            if (full) { o[3] = (uint16_t)(w1 >> 16); o[4] = (uint16_t)w2; o[5] = (uint16_t)(w2 >> 16); }
        }
    }
}

// Interior rows: the strip row's RGB bytes (62 producing lanes x 4 px) are
// regrouped through a per-wave LDS slice so that the global stores are
// 16-byte, lane-contiguous and write each cache line once (the direct
// 24- / 12-byte-per-lane stores of store4 touch every line twice).  seg_px =
// index of the strip's first pixel (y * w + strip start).  u16 output: 1488
// bytes = 64 + 29 lanes x 16 B; u8 output (w % 4 == 0 only): 744 bytes =
// 46 lanes x 16 B + 8 B.  LDS accesses of one wave execute in order; the
// wave-scope fences only keep the compiler from reordering them.
__device__ __forceinline__ void lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One 16-byte chunk of a staged row at byte off of the segment, clipped to
// nbytes (a multiple of 4).  Whole chunks are non-temporal stores: the RGB
// output is streamed once and never re-read by this kernel, and keeping it
// out of the caches' allocation policy lets the writes reach HBM at close to
// the write-only streaming rate (measured: 12 MP MG frame 27.8 -> 20.2 us).
__device__ __forceinline__ void store_chunk(uint8_t* base, uint32_t off, uint4 v, uint32_t nbytes) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    if (off + 16u <= nbytes) {
        const u32x4 t = {v.x, v.y, v.z, v.w};
        __builtin_nontemporal_store(t, (u32x4*)(base + off));
    } else if (off < nbytes) {
        uint32_t* p = (uint32_t*)(base + off);
        p[0] = v.x;
        if (off + 8u <= nbytes) p[1] = v.y;
        if (off + 12u <= nbytes) p[2] = v.z;
    }
}

// Strip geometry (strip_stride / strip_count): the wave's computed columns
// start at xs (lane run x = xs - 4 + 4 * lane, lanes 1..62 produce), and in
// row y it writes columns [lo, hi) (seg_cols).
struct Strip {
    int xs;
    bool first, last;
};

__device__ __forceinline__ bool aligned_strips(const FastParams& f) { return f.out16 && f.aligned; }

__device__ __forceinline__ Strip strip_of(const FastParams& f, int sx) {
    return Strip{sx * strip_stride(aligned_strips(f)), sx == 0, sx == f.nsx - 1};
}

// s(y): aligned strips, the shift in {0, 8, 16, 24} that puts the row's strip
// boundaries on 64-byte boundaries (48 * s/8 == -(out + 6 * y * w) mod 64
// needs that address to be a multiple of 16; otherwise 0).  Only the low
// bits of the address matter, so 32-bit wrap-around is harmless.
__device__ __forceinline__ int row_shift(const FastParams& f, int y) {
    if (!aligned_strips(f)) return 0;
    const uint32_t a = (uint32_t)(uintptr_t)f.out + 6u * (uint32_t)y * (uint32_t)f.w;
    if (a & 15u) return 0;
    const uint32_t t = ((64u - (a & 63u)) & 63u) >> 4;  // -a / 16 (mod 4)
    return 8 * (int)((3u * t) & 3u);                     // 3 * 3 == 1 (mod 4)
}

__device__ __forceinline__ void seg_cols(const FastParams& f, const Strip& sp, int y, int& lo, int& hi) {
    const int sh = row_shift(f, y);
    lo = sp.first ? 0 : sp.xs + sh;
    hi = sp.last ? f.w : imin(sp.xs + strip_stride(aligned_strips(f)) + sh, f.w);
}

// This lane's 4 pixels at x lie in the row's write window (windows start at
// multiples of 8 columns, so a lane run is wholly in or out).
__device__ __forceinline__ bool lane_writes(const FastParams& f, const Strip& sp, int x, int y) {
    const int lane = threadIdx.x & 63;
    int lo, hi;
    seg_cols(f, sp, y, lo, hi);
    return lane >= 1 && lane <= 62 && x >= lo && x < hi;
}

// The staged row holds the producing lanes' 248 pixels from column xs (lane
// l at pixel 4 * (l - 1)); columns [c0, c1) of it (relative to xs) are
// written at row pixel index row_px = y * w + xs.  c0 * (bytes per pixel) is
// a multiple of 16 (u16: c0 in {0, 8, 16, 24}; u8: c0 == 0), so the window
// is read from LDS in whole 16-byte chunks, and no chunk read lies past the
// staged 1488 bytes by more than one chunk (the slice holds 1536).
__device__ __forceinline__ void store_row_staged(const FastParams& f, uint32_t* stage, int lane, uint32_t row_px,
                                                 int c0, int c1, const int r[4], const int g[4], const int b[4],
                                                 bool wr) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    // window arithmetic is wave-uniform: pinned to scalar registers before
    // the lane-divergent staging writes
    const uint32_t e = __builtin_amdgcn_readfirstlane(f.out16 ? 6u : 3u);
    const uint32_t b0 = __builtin_amdgcn_readfirstlane(e * (uint32_t)c0);
    const uint32_t nbytes = __builtin_amdgcn_readfirstlane(e * (uint32_t)(c1 - c0));
    uint8_t* base = f.out + (uint64_t)e * (row_px + (uint32_t)c0);
    if (f.out16) {
        if (wr) {
            uint2* p = (uint2*)(stage + 6 * (lane - 1));
            p[0] = make_uint2((uint32_t)r[0] | ((uint32_t)g[0] << 16), (uint32_t)b[0] | ((uint32_t)r[1] << 16));
            p[1] = make_uint2((uint32_t)g[1] | ((uint32_t)b[1] << 16), (uint32_t)r[2] | ((uint32_t)g[2] << 16));
            p[2] = make_uint2((uint32_t)b[2] | ((uint32_t)r[3] << 16), (uint32_t)g[3] | ((uint32_t)b[3] << 16));
        }
    } else {
        if (wr) {
            uint32_t* p = stage + 3 * (lane - 1);
            p[0] = (uint32_t)r[0] | ((uint32_t)g[0] << 8) | ((uint32_t)b[0] << 16) | ((uint32_t)r[1] << 24);
            p[1] = (uint32_t)g[1] | ((uint32_t)b[1] << 8) | ((uint32_t)r[2] << 16) | ((uint32_t)g[2] << 24);
            p[2] = (uint32_t)b[2] | ((uint32_t)r[3] << 8) | ((uint32_t)g[3] << 16) | ((uint32_t)b[3] << 24);
        }
    }
    const uint32_t o0 = 16u * (uint32_t)lane, o1 = 1024u + o0;
    const bool two = lane < 32 && o1 < nbytes;
    lds_order();
    const uint4 v0 = ((const uint4*)stage)[(b0 >> 4) + (uint32_t)lane];
    uint4 v1 = make_uint4(0u, 0u, 0u, 0u);
    if (two) v1 = ((const uint4*)stage)[(b0 >> 4) + 64u + (uint32_t)lane];
    lds_order();
    if ((nbytes & 15u) == 0u) {  // whole chunks (every interior u16 window)
        if (o0 < nbytes) __builtin_nontemporal_store((u32x4){v0.x, v0.y, v0.z, v0.w}, (u32x4*)(base + o0));
        if (two) __builtin_nontemporal_store((u32x4){v1.x, v1.y, v1.z, v1.w}, (u32x4*)(base + o1));
    } else {
        store_chunk(base, o0, v0, nbytes);
        if (two) store_chunk(base, o1, v1, nbytes);
    }
}

// store_row_staged for u16 RGB rows whose window is whole 16-byte chunks
// (interior rows of aligned strips: 1344 bytes): lanes 0 and 63 stage their
// unused pixels past the 1488 staged bytes instead of branching around the
// LDS writes, and both chunk reads are unconditional (the second may read
// past the staged row, inside the caller's LDS).
__device__ __forceinline__ void store_row_u16_whole(const FastParams& f, uint32_t* stage, int lane, uint32_t row_px,
                                                    int c0, int c1, const int r[4], const int g[4],
                                                    const int b[4]) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const uint32_t b0 = __builtin_amdgcn_readfirstlane(6u * (uint32_t)c0);
    const uint32_t nbytes = __builtin_amdgcn_readfirstlane(6u * (uint32_t)(c1 - c0));
    uint8_t* base = f.out + 6ull * (row_px + (uint32_t)c0);
    const uint32_t slot = (uint32_t)(lane - 1) < 62u ? 6u * (uint32_t)(lane - 1) : 372u;
    uint2* p = (uint2*)(stage + slot);
    p[0] = make_uint2((uint32_t)r[0] | ((uint32_t)g[0] << 16), (uint32_t)b[0] | ((uint32_t)r[1] << 16));
    p[1] = make_uint2((uint32_t)g[1] | ((uint32_t)b[1] << 16), (uint32_t)r[2] | ((uint32_t)g[2] << 16));
    p[2] = make_uint2((uint32_t)b[2] | ((uint32_t)r[3] << 16), (uint32_t)g[3] | ((uint32_t)b[3] << 16));
    const uint32_t o0 = 16u * (uint32_t)lane, o1 = 1024u + o0;
    lds_order();
    const uint4 v0 = ((const uint4*)stage)[(b0 >> 4) + (uint32_t)lane];
    const uint4 v1 = ((const uint4*)stage)[(b0 >> 4) + 64u + (uint32_t)lane];
    lds_order();
    if (o0 < nbytes) __builtin_nontemporal_store((u32x4){v0.x, v0.y, v0.z, v0.w}, (u32x4*)(base + o0));
    if (o1 < nbytes) __builtin_nontemporal_store((u32x4){v1.x, v1.y, v1.z, v1.w}, (u32x4*)(base + o1));
}

// MG colour difference of the two R/B sites a lane owns in one row.
// P points at 5 window rows centred on the row (P[0] two rows up).
// Halo columns come from the neighbouring lanes.
template <int PAT, int MODE, bool CAREFUL>
__device__ __forceinline__ void d_row(const int (*P)[4], int py, int row_ok, const bool col_ok[4], int D[3]) {
    const int pc = rb_col_parity<PAT>(py);
    const int* c = P[2];
    // row values at lane columns -2..5
    int v[8];
    v[0] = from_left(c[2]);
    v[1] = from_left(c[3]);
    v[2] = c[0]; v[3] = c[1]; v[4] = c[2]; v[5] = c[3];
    v[6] = from_right(c[0]);
    v[7] = from_right(c[1]);
    int own[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int cj = pc + 2 * j;  // lane column of the site
        const int cc = v[cj + 2];
        const int nr = P[1][cj] + P[3][cj] + v[cj + 1] + v[cj + 3];
        const int fr = P[0][cj] + P[4][cj] + v[cj] + v[cj + 4];
        const int gn = 4 * cc + 2 * nr - fr;
        int d;
        if (MODE == MODE_MG_EXACT) d = 8 * cc - gn;
        else d = cc - clamp0((gn + 4) >> 3, 255);
        if (CAREFUL) {
            const int sentinel = (MODE == MODE_MG_EXACT) ? 8 * cc + 8 : cc;
            d = (row_ok && col_ok[cj]) ? d : sentinel;
        }
        own[j] = d;
    }
    if (pc == 0) { D[0] = own[0]; D[1] = own[1]; D[2] = from_right(own[0]); }
    else { D[0] = from_left(own[1]); D[1] = own[0]; D[2] = own[1]; }
}

// D value of the R/B site at lane column cr (-1..4) of a D row of parity py.
template <int PAT>
__device__ __forceinline__ int dat(const int D[3], int py, int cr) {
    return D[(cr + rb_col_parity<PAT>(py)) >> 1];
}

// One output pixel of the step from the window: row a (0/1) of the quad row,
// lane column k (0..3).  D rows: Dm = row above, D0 = own row, Dp = row below.
template <int PAT, int MODE>
__device__ __forceinline__ void mg_px(int a, int k, int c, const int* Dm, const int* D0, const int* Dp, int ch,
                                      int& r, int& g, int& b) {
    const int s = role<PAT>(a, k & 1);
    const int pm = (a + 1) & 1, p0 = a & 1, pp = (a + 1) & 1;
    if (s != 1) {
        const int dc = dat<PAT>(D0, p0, k);
        const int ds = dat<PAT>(Dm, pm, k - 1) + dat<PAT>(Dm, pm, k + 1) + dat<PAT>(Dp, pp, k - 1) +
                       dat<PAT>(Dp, pp, k + 1);
        int gg, oth;
        if (MODE == MODE_MG_EXACT) {
            const int t = 8 * c - dc + 4;           // 8 * G + 4
            gg = t >> 3;
            oth = (4 * t + ds) >> 5;                 // (32c - 4D + sum D + 16) >> 5
        } else {
            gg = c - dc;
            oth = gg + tdiv4(ds + 2);
        }
        gg = clamp0(gg, ch);
        oth = clamp0(oth, ch);
        if (s == 0) { r = c; g = gg; b = oth; } else { b = c; g = gg; r = oth; }
    } else {
        const int hs = dat<PAT>(D0, p0, k - 1) + dat<PAT>(D0, p0, k + 1);
        const int vs = dat<PAT>(Dm, pm, k) + dat<PAT>(Dp, pp, k);
        int hz, vt;
        if (MODE == MODE_MG_EXACT) {
            const int u = 16 * c + 8;
            hz = (u + hs) >> 4;
            vt = (u + vs) >> 4;
        } else {
            hz = c + tdiv2(hs + 1);
            vt = c + tdiv2(vs + 1);
        }
        hz = clamp0(hz, ch);
        vt = clamp0(vt, ch);
        g = c;
        if (a == r_row_of(PAT)) { r = hz; b = vt; } else { r = vt; b = hz; }
    }
}

// Raw bytes of the lane run in row y.  Interior bands: every row and lane
// is inside the frame, so no masking; their 8-byte windows end
// before the stream does because the final row is never in their window.
// VRIM: the band's rows reach the top or bottom rim (rows outside the frame,
// the stream's final row); CAREFUL without VRIM: only lanes outside the
// frame's columns, with every row inside and followed by another.
template <int FMT, bool CAREFUL, bool ODD, bool VRIM = CAREFUL>
__device__ __forceinline__ Raw load_row(const FastParams& f, int y, uint32_t xu, bool lane_ok, uint32_t rs,
                                        uint32_t lo) {
    constexpr bool packed = FMT == FMT_P12 || FMT == FMT_P10 || FMT == FMT_G10;
    const bool row_ok = !VRIM || (y >= 0 && y < f.h);
    if (ODD) {
        // the last row is decoded pixel by pixel (its final group may be
        // cut short); every other row has >= 11 stream bytes past the run
        const bool ok = row_ok && lane_ok && !(VRIM && y == f.h - 1);
        return load_raw_any<FMT>(f, ok ? run_offset<FMT>((uint32_t)f.w, xu, (uint32_t)y) : 0u);
    }
    const uint32_t off = (row_ok ? (uint32_t)y : 0u) * rs + lo;
    if (VRIM && packed) return load_raw_exact<FMT>(f, off, row_ok && lane_ok);
    return load_raw<FMT>(f, off, !CAREFUL || (row_ok && lane_ok));
}

// sqrt(t) rounded to nearest for t == 0 or 2^-96 <= t < inf: the hardware
// estimate (within 1 ulp) replaced by its lower or upper neighbour when the
// fused residual t - neighbour * estimate says that one is nearer (the
// correction the compiler's correctly rounded sqrtf applies, without the
// rescaling it needs only below 2^-96).
__device__ __forceinline__ float sqrt_rn(float t) {
    const float s = __builtin_amdgcn_sqrtf(t);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u);
    const float sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float lo = fmaf(-sm, s, t) <= 0.0f ? sm : s;
    return fmaf(-sp, s, t) > 0.0f ? sp : lo;
}

// a / b rounded to nearest from y = 1/b rounded to nearest: two fused
// remainder corrections of a * y (Markstein); exact while no intermediate
// is subnormal or overflows.
__device__ __forceinline__ float div_rn(float a, float b, float y) {
#pragma clang fp contract(off)
    const float q0 = a * y;
    const float q1 = fmaf(fmaf(-b, q0, a), y, q0);
    return fmaf(fmaf(-b, q1, a), y, q1);
}

// Lens shading of one decoded, black-levelled sample at (x, y) of CFA parity
// `parity`: the reference's operations (polynomial.cpp:30-47) in its order
// and contractions -- r = sqrt(fma(dx, dx, dy * dy)) / max_r, gain =
// fma(a6, r^6, fma(a4, r^4, fma(a2, r^2, 1))), v' = clamp(int(fma(v, gain,
// 0.5)), 0, max) -- with sqrt_rn / div_rn for the rounded sqrt and quotient
// and the clamp as a median (the host takes this path only for frames where
// those are exact: LensParams::fused_ok).
__device__ __forceinline__ int lens_px(const FastParams& f, int v, int x, float dyy, int parity) {
#pragma clang fp contract(off)
    const float dx = (float)x - f.lens.cx;
    const float r = div_rn(sqrt_rn(fmaf(dx, dx, dyy)), f.lens.max_r, f.lens.inv_max_r);
    const float r2 = r * r, r4 = r2 * r2, r6 = r4 * r2;
    const float* a = f.lens.pc[parity];
    const float gain = fmaf(a[2], r6, fmaf(a[1], r4, fmaf(a[0], r2, 1.0f)));
    return (int)__builtin_amdgcn_fmed3f(fmaf((float)v, gain, 0.5f), 0.0f, (float)f.maxv);
}

template <int FMT, bool CAREFUL, bool ODD, int YODD, bool VRIM, bool LENS>
__device__ __forceinline__ void decode_row_core(const FastParams& f, int x, int y, Raw raw, bool lane_ok,
                                                uint32_t xu, int P[4]);

// Decoded, black-levelled samples of row y (parity YODD) from its raw bytes;
// LENS: then the lens-shading gain (the fused Bayer prologue).
template <int FMT, bool CAREFUL, bool ODD, int YODD, bool VRIM = CAREFUL, bool LENS = false>
__device__ __forceinline__ void decode_row(const FastParams& f, int x, int y, Raw raw, bool lane_ok, uint32_t xu,
                                           int P[4]) {
    decode_row_core<FMT, CAREFUL, ODD, YODD, VRIM, LENS>(f, x, y, raw, lane_ok, xu, P);
    if (LENS) {
#pragma clang fp contract(off)
        const float dy = (float)y - f.lens.cy;
        const float dyy = dy * dy;
#pragma unroll
        for (int k = 0; k < 4; ++k) P[k] = lens_px(f, P[k], x + k, dyy, YODD * 2 + (k & 1));
    }
}

template <int FMT, bool CAREFUL, bool ODD, int YODD, bool VRIM, bool LENS>
__device__ __forceinline__ void decode_row_core(const FastParams& f, int x, int y, Raw raw, bool lane_ok,
                                                uint32_t xu, int P[4]) {
    const int o0 = f.bl_i[YODD * 2], o1 = f.bl_i[YODD * 2 + 1];
    const bool row_ok = !VRIM || (y >= 0 && y < f.h);
    uint32_t v[4];
    if (ODD && VRIM && y == f.h - 1) {
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = (lane_ok && x + k < f.w) ? raw_pixel<FMT>(f, xu + k, (uint32_t)y) : 0u;
    } else if (FMT == FMT_P12 || FMT == FMT_G10) {
        if (FMT == FMT_P12) p12_samples(raw, o0, o1, v);  // black level included
        else g10_samples(raw, o0, o1, v);
        const bool ok = !CAREFUL || (row_ok && lane_ok);
#pragma unroll
        for (int k = 0; k < 4; ++k) P[k] = ok ? (int)v[k] : 0;
        return;
    } else {
        raw_values<FMT>(raw, ODD && YODD, v);  // w % 4 == 2: odd rows start at sub-pixel 2
    }
    finish4<FMT, CAREFUL, LENS>(f, v, row_ok && lane_ok, o0, o1, P);
}

// Interior site columns of the lane run (MG green kernel) and the columns the
// colour-difference pass reaches.
__device__ __forceinline__ void lane_columns(int x, int w, bool col_ok[4], bool col_core[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        col_ok[k] = (x + k >= 2) && (x + k < w - 2);
        col_core[k] = (x + k >= 1) && (x + k < w - 1);
    }
}

// Last gamma table entry the demosaic's output can index: 255 for u8 RGB,
// the bit depth's maximum for u16 RGB (every output is clamped to it).
__device__ __forceinline__ uint32_t lut_last(const FastParams& f) {
    return f.out16 ? (uint32_t)f.maxv : 255u;
}

// One step: output rows yA, yA + 1 (yA even) of the lane run at x.  Pw holds
// the decoded rows yA-1 .. yA+4; Dw the MG colour differences of rows yA-1,
// yA on entry, and the step forms rows yA+1, yA+2 into Dw[2], Dw[3] (DGIVEN:
// all four are given, and of Pw only the rows the step reads -- yA, yA+1 for
// MG, yA-1 .. yA+2 for HQLI -- need to hold values).
template <int FMT, int PAT, int MODE, bool CAREFUL, bool ODD, bool EPI, bool DGIVEN = false, int OUT = 0>
__device__ __forceinline__ void band_step(const FastParams& f, const Strip& sp, int x, int yA, uint32_t* stage,
                                          const int (*Pw)[4], int (*Dw)[3], const bool col_ok[4],
                                          const bool col_core[4], const uint16_t* lut) {
    const int w = f.w, h = f.h;
    const int ch = MODE == MODE_MG_INT8 ? 255 : f.clamp_hi;
    if (MODE != MODE_HQLI && !DGIVEN) {
        d_row<PAT, MODE, CAREFUL>(&Pw[0], 1, (yA + 1 >= 2) && (yA + 1 < h - 2), col_ok, Dw[2]);
        d_row<PAT, MODE, CAREFUL>(&Pw[1], 0, (yA + 2 >= 2) && (yA + 2 < h - 2), col_ok, Dw[3]);
    }
    int R[2][4], G[2][4], B[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int* p = Pw[1 + a];
        const int* pu = Pw[a];
        const int* pd = Pw[2 + a];
        int hl = 0, hr = 0, ul = 0, ur = 0, dl = 0, dr = 0;
        if (MODE == MODE_HQLI) {
            hl = from_left(p[3]); hr = from_right(p[0]);
            ul = from_left(pu[3]); ur = from_right(pu[0]);
            dl = from_left(pd[3]); dr = from_right(pd[0]);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = p[k];
            int r, g, b;
            if (MODE == MODE_HQLI) {
                const int s = role<PAT>(a, k & 1);
                const int le = k > 0 ? p[k - 1] : hl, ri = k < 3 ? p[k + 1] : hr;
                const int up = pu[k], dn = pd[k];
                if (s != 1) {
                    const int ul_ = k > 0 ? pu[k - 1] : ul, ur_ = k < 3 ? pu[k + 1] : ur;
                    const int dl_ = k > 0 ? pd[k - 1] : dl, dr_ = k < 3 ? pd[k + 1] : dr;
                    const int gg = (le + ri + up + dn + 2) >> 2;
                    const int oth = (ul_ + ur_ + dl_ + dr_ + 2) >> 2;
                    if (s == 0) { r = c; g = gg; b = oth; } else { b = c; g = gg; r = oth; }
                } else {
                    const int hz = (le + ri + 1) >> 1, vt = (up + dn + 1) >> 1;
                    g = c;
                    if (a == r_row_of(PAT)) { r = hz; b = vt; } else { r = vt; b = hz; }
                }
                r = imin(r, ch); g = imin(g, ch); b = imin(b, ch);
            } else {
                mg_px<PAT, MODE>(a, k, c, Dw[a], Dw[a + 1], Dw[a + 2], ch, r, g, b);
                if (CAREFUL) {
                    // positions the colour-difference pass never reaches keep R = B = 0
                    const bool core = col_core[k] && (yA + a >= 1) && (yA + a < h - 1);
                    r = core ? r : 0;
                    b = core ? b : 0;
                }
            }
            R[a][k] = r; G[a][k] = g; B[a][k] = b;
        }
    }
    // RGB epilogue (white balance, colour matrix, gamma) on every pixel
    // computed, before the border fill copies any of them
    if (EPI) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int k = 0; k < 4; ++k) epi_px(f.epi, R[a][k], G[a][k], B[a][k], LutRead{lut, lut_last(f)});
    }
    // w % 4 == 2: column w-3 of the right-edge column fill lives in the
    // lane to the left (DPP while every lane is active)
    int LR[2] = {0, 0}, LG[2] = {0, 0}, LB[2] = {0, 0};
    if (ODD && CAREFUL) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            LR[a] = from_left(R[a][3]);
            LG[a] = from_left(G[a][3]);
            LB[a] = from_left(B[a][3]);
        }
    }

#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int y = yA + a;
        const int lane = threadIdx.x & 63;
        if (!CAREFUL) {
            // u16 RGB in aligned strips (OUT 1: the single-frame kernel on a
            // deeper-than-8-bit frame): every interior window is whole chunks
            if (OUT == 1) {
                int lo, hi;
                seg_cols(f, sp, y, lo, hi);
                store_row_u16_whole(f, stage, lane, (uint32_t)y * (uint32_t)w + (uint32_t)sp.xs, lo - sp.xs,
                                    hi - sp.xs, R[a], G[a], B[a]);
                continue;
            }
            // staged rows need 4-byte aligned runs: always for u16, for u8 when w % 4 == 0
            if (!ODD || f.out16) {
                int lo, hi;
                seg_cols(f, sp, y, lo, hi);
                store_row_staged(f, stage, lane, (uint32_t)y * (uint32_t)w + (uint32_t)sp.xs, lo - sp.xs,
                                 hi - sp.xs, R[a], G[a], B[a], lane >= 1 && lane <= 62);
            } else if (lane_writes(f, sp, x, y)) {
                store4<ODD>(f, (uint32_t)x, (uint32_t)y, R[a], G[a], B[a]);
            }
            continue;
        }
        if (y >= h || y < 2 || y >= h - 2) continue;  // the same for every lane
        int r[4], g[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) { r[k] = R[a][k]; g[k] = G[a][k]; b[k] = B[a][k]; }
        if (MODE == MODE_HQLI) {
            // pre-fill value of a border column: never written -> 0
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (!col_ok[k]) { r[k] = 0; g[k] = 0; b[k] = 0; }
        }
        if (y == 2) {
            if (lane_writes(f, sp, x, 0)) store4<ODD>(f, x, 0, r, g, b);
            if (lane_writes(f, sp, x, 1)) store4<ODD>(f, x, 1, r, g, b);
        }
        if (y == h - 3) {
            if (lane_writes(f, sp, x, h - 2)) store4<ODD>(f, x, h - 2, r, g, b);
            if (lane_writes(f, sp, x, h - 1)) store4<ODD>(f, x, h - 1, r, g, b);
        }
        // column fill of the middle rows: columns 0-1 <- 2, w-2..w-1 <- w-3
        if (x == 0) { r[0] = r[1] = R[a][2]; g[0] = g[1] = G[a][2]; b[0] = b[1] = B[a][2]; }
        if (!ODD && x == w - 4) { r[2] = r[3] = R[a][1]; g[2] = g[3] = G[a][1]; b[2] = b[3] = B[a][1]; }
        if (ODD && x == w - 2) { r[0] = r[1] = LR[a]; g[0] = g[1] = LG[a]; b[0] = b[1] = LB[a]; }
        if (!ODD || f.out16) {
            // staged like the interior rows; the window ends at the frame's right edge
            int lo, hi;
            seg_cols(f, sp, y, lo, hi);
            store_row_staged(f, stage, lane, (uint32_t)y * (uint32_t)w + (uint32_t)sp.xs, lo - sp.xs, hi - sp.xs,
                             r, g, b, lane >= 1 && lane <= 62 && x < w);
        } else if (lane_writes(f, sp, x, y)) {
            store4<ODD>(f, x, y, r, g, b);
        }
    }
}

// One band of TH rows from y0 (even) for the lane run at x.  ODD: w % 4 == 2.
// Every load of the band is issued up front.
template <int FMT, int PAT, int MODE, bool CAREFUL, int TH, bool ODD, bool EPI = false, bool VRIM = CAREFUL,
          bool LENS = false>
__device__ __forceinline__ void run_band(const FastParams& f, const Strip& sp, int y0, uint32_t* stage,
                                         const uint16_t* lut = nullptr) {
    constexpr int NR = TH + 6;  // raw rows y0-3 .. y0+TH+2
    const int x = sp.xs - 4 + 4 * (int)(threadIdx.x & 63);
    const bool lane_ok = x >= 0 && x < f.w;
    const uint32_t xu = lane_ok ? (uint32_t)x : 0u;
    bool col_ok[4], col_core[4];
    lane_columns(x, f.w, col_ok, col_core);
    const uint32_t rs = row_stride<FMT>((uint32_t)f.w), lo = lane_offset<FMT>(xu);
    Raw raw[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) raw[i] = load_row<FMT, CAREFUL, ODD, VRIM>(f, y0 - 3 + i, xu, lane_ok, rs, lo);

    int P[NR][4];  // row y0 - 3 + i; y0 even, so row i is odd when i is even
    auto decode = [&](int i) {
        if (i & 1) decode_row<FMT, CAREFUL, ODD, 0, VRIM, LENS>(f, x, y0 - 3 + i, raw[i], lane_ok, xu, P[i]);
        else decode_row<FMT, CAREFUL, ODD, 1, VRIM, LENS>(f, x, y0 - 3 + i, raw[i], lane_ok, xu, P[i]);
    };
#pragma unroll
    for (int i = 0; i < 6; ++i) decode(i);
    int D[TH + 2][3];  // MG colour differences, rows y0-1 .. y0+TH
    if (MODE != MODE_HQLI) {
        d_row<PAT, MODE, CAREFUL>(&P[0], 1, (y0 - 1 >= 2) && (y0 - 1 < f.h - 2), col_ok, D[0]);
        d_row<PAT, MODE, CAREFUL>(&P[1], 0, (y0 >= 2) && (y0 < f.h - 2), col_ok, D[1]);
    }
#pragma unroll
    for (int st = 0; st < TH / 2; ++st) {
        decode(2 * st + 6);
        decode(2 * st + 7);
        band_step<FMT, PAT, MODE, CAREFUL, ODD, EPI>(f, sp, x, y0 + 2 * st, stage, &P[2 * st + 2], &D[2 * st],
                                                     col_ok, col_core, lut);
    }
}


// ------------------------------------------------------ single-frame groups
//
// The single-frame kernel runs workgroups of GB waves stacked down one strip:
// wave W of group g computes rows yg + TH W .. yg + TH W + TH - 1 (yg =
// g TH GB).  The group decodes each of its TH GB + 6 raw rows once and forms
// each of its TH GB + 2 MG colour-difference rows once, sharing both through
// LDS, where a wave on its own decodes TH + 6 rows and forms TH + 2 for TH
// output rows:
//   phase 1  wave W decodes rows W, W + GB, ... of yg - 3 .. yg + TH GB + 2
//            into LDS as 16-bit sample pairs (P0, P2) and (P1, P3);
//   phase 2  (MG) wave W forms colour-difference rows W, W + GB, ... of
//            yg - 1 .. yg + TH GB from five LDS rows each, its two R/B sites
//            per lane into LDS;
//   phase 3  every wave reads its own rows and the colour differences around
//            them and runs band_step (interpolation, border fill, epilogue,
//            staged stores) as the per-wave form does.
// Groups whose rows or columns touch the frame rim run the same phases with
// the careful forms (masked loads, sentinels, border fill).

__device__ __forceinline__ void lds_sync() { __syncthreads(); }

// LDS layout of a group (bytes): decoded rows (512 B: x[64] then y[64]),
// colour-difference rows (512 B: own site 0 [64] then own site 1 [64]), and,
// for 2-row waves, separate 1536-byte staging slices.  With 4-row waves a
// wave's staging slice aliases LDS no other wave reads in phase 3: its own
// decoded rows under MG (read into registers first), the colour-difference
// rows under HQLI (which has none).  The staged-row reads of
// store_row_u16_whole reach up to 2,191 bytes past a slice's start (a
// 144-byte window offset plus two 1 KB chunk rows), so the separate slices
// are followed by 768 bytes that keep the last wave's reads inside the group.
template <int TH, int GB>
struct GroupLayout {
    static constexpr int kRows = TH * GB;
    static constexpr int kPR = kRows + 6;
    static constexpr int kDR = kRows + 2;
    static constexpr int kDOff = kPR * 512;
    static constexpr int kSOff = kDOff + kDR * 512;
    static constexpr bool kSepStage = TH < 4;
    static constexpr int kBytes = kSOff + (kSepStage ? GB * 1536 + 768 : 0);
};

__device__ __forceinline__ void unpack4(uint32_t X, uint32_t Y, int p[4]) {
    p[0] = (int)(X & 0xFFFFu);
    p[2] = (int)(X >> 16);
    p[1] = (int)(Y & 0xFFFFu);
    p[3] = (int)(Y >> 16);
}

// D[3] of a row (as d_row forms it) from the lane's own two sites
template <int PAT>
__device__ __forceinline__ void d_from_own(int py, int own0, int own1, int D[3]) {
    if (rb_col_parity<PAT>(py) == 0) {
        D[0] = own0; D[1] = own1; D[2] = from_right(own0);
    } else {
        D[0] = from_left(own1); D[1] = own0; D[2] = own1;
    }
}

// One wave W (compile time) of a group whose rows lie inside the frame's
// vertical rim.  HC: the group's strip reaches the left or right rim, so
// loads and samples of lanes outside the frame are masked, colour
// differences take the sentinels of the columns the pass never reaches,
// and the output takes the border column fill (band_step's careful form).
template <int FMT, int PAT, int MODE, int TH, int GB, bool ODD, bool EPI, bool LENS, int W, bool HC = false>
__device__ __forceinline__ void group_wave(const FastParams& f, const Strip& sp, int yg, uint8_t* L,
                                           const uint16_t* lut) {
    using G = GroupLayout<TH, GB>;
    const int lane = threadIdx.x & 63;
    const int x = sp.xs - 4 + 4 * lane;
    const bool lane_ok = x >= 0 && x < f.w;
    const uint32_t xu = lane_ok ? (uint32_t)x : 0u;
    bool col_ok[4], col_core[4];
    lane_columns(x, f.w, col_ok, col_core);
    const uint32_t rs = row_stride<FMT>((uint32_t)f.w), lo = lane_offset<FMT>(xu);
    uint32_t* Lw = (uint32_t*)L;  // dword view
    // phase 1: decoded rows i = W + GB k (row yg - 3 + i, parity (W + 1) & 1)
    constexpr int NI = (G::kPR - W + GB - 1) / GB;
    constexpr int YODD = (W + 1) & 1;
    {
        Raw raw[NI];
#pragma unroll
        for (int k = 0; k < NI; ++k)
            raw[k] = load_row<FMT, HC, ODD, false>(f, yg - 3 + W + GB * k, xu, lane_ok, rs, lo);
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int i = W + GB * k;
            int P[4];
            decode_row<FMT, HC, ODD, YODD, false, LENS>(f, x, yg - 3 + i, raw[k], lane_ok, xu, P);
            Lw[i * 128 + lane] = (uint32_t)P[0] | ((uint32_t)P[2] << 16);
            Lw[i * 128 + 64 + lane] = (uint32_t)P[1] | ((uint32_t)P[3] << 16);
        }
    }
    lds_sync();
    // phase 2 (MG): colour-difference rows j = W + GB k (row yg - 1 + j,
    // parity 1 - (W & 1)) from decoded rows j .. j + 4
    if (MODE != MODE_HQLI) {
        constexpr int NJ = (G::kDR - W + GB - 1) / GB;
        constexpr int PY = 1 - (W & 1);
        constexpr int PC = rb_col_parity<PAT>(PY);
#pragma unroll
        for (int k = 0; k < NJ; ++k) {
            const int j = W + GB * k;
            int P5[5][4];
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                if (r == 2) {
                    unpack4(Lw[(j + 2) * 128 + lane], Lw[(j + 2) * 128 + 64 + lane], P5[2]);
                } else {  // only the R/B-site columns of this row's parity are read
                    const uint32_t v = Lw[(j + r) * 128 + (PC == 0 ? 0 : 64) + lane];
                    P5[r][PC] = (int)(v & 0xFFFFu);
                    P5[r][PC + 2] = (int)(v >> 16);
                    P5[r][1 - PC] = 0;
                    P5[r][3 - PC] = 0;
                }
            }
            int D[3];
            d_row<PAT, MODE, HC>(P5, PY, 1, col_ok, D);
            Lw[G::kDOff / 4 + j * 128 + lane] = (uint32_t)(PC == 0 ? D[0] : D[1]);
            Lw[G::kDOff / 4 + j * 128 + 64 + lane] = (uint32_t)(PC == 0 ? D[1] : D[2]);
        }
    }
    lds_sync();
    // phase 3: output rows yg + TH W .. yg + TH W + TH - 1 (inside the frame:
    // interior groups end 3 rows above its last row)
    const int yw = yg + TH * W;
    constexpr int pw = TH * W + 3;  // decoded row of yw
    uint32_t cx[TH], cy[TH];
    if (MODE != MODE_HQLI) {
#pragma unroll
        for (int r = 0; r < TH; ++r) {
            cx[r] = Lw[(pw + r) * 128 + lane];
            cy[r] = Lw[(pw + r) * 128 + 64 + lane];
        }
    }
    uint32_t* stage;
    if (G::kSepStage) stage = Lw + G::kSOff / 4 + W * 384;
    else if (MODE != MODE_HQLI) stage = Lw + pw * 128;  // own decoded rows, read above
    else stage = Lw + G::kDOff / 4 + W * 384;            // HQLI: no colour-difference rows
    // MG into u16 RGB without the epilogue: interpolation and whole-chunk
    // staged stores inline; every other form through band_step (8-bit
    // frames write u16 RGB only on request, as batch frames do)
    constexpr bool lean_form = MODE != MODE_HQLI && !EPI && !HC;
    const bool lean = lean_form && (FMT != FMT_U8 || f.out16);
    const int ch = MODE == MODE_MG_INT8 ? 255 : f.clamp_hi;
#pragma unroll
    for (int s = 0; s < TH / 2; ++s) {
        const int yA = yw + 2 * s;
        int Pw[6][4];
        if (MODE == MODE_HQLI) {
#pragma unroll
            for (int r = 0; r < 4; ++r)  // rows yA - 1 .. yA + 2
                unpack4(Lw[(pw + 2 * s - 1 + r) * 128 + lane], Lw[(pw + 2 * s - 1 + r) * 128 + 64 + lane], Pw[r]);
        } else {
            unpack4(cx[2 * s], cy[2 * s], Pw[1]);
            unpack4(cx[2 * s + 1], cy[2 * s + 1], Pw[2]);
        }
        int Dw[4][3];
        if (MODE != MODE_HQLI) {
            const int jb = TH * W + 2 * s;  // colour-difference row of yA - 1
#pragma unroll
            for (int t = 0; t < 4; ++t)
                d_from_own<PAT>((t + 1) & 1, (int)Lw[G::kDOff / 4 + (jb + t) * 128 + lane],
                                (int)Lw[G::kDOff / 4 + (jb + t) * 128 + 64 + lane], Dw[t]);
        }
        if (lean_form && lean) {
            int R[2][4], Gc[2][4], B[2][4];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    mg_px<PAT, MODE>(a, k, Pw[1 + a][k], Dw[a], Dw[a + 1], Dw[a + 2], ch, R[a][k], Gc[a][k], B[a][k]);
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int y = yA + a;
                int l0, h0;
                seg_cols(f, sp, y, l0, h0);
                store_row_u16_whole(f, stage, lane, (uint32_t)y * (uint32_t)f.w + (uint32_t)sp.xs, l0 - sp.xs,
                                    h0 - sp.xs, R[a], Gc[a], B[a]);
            }
        } else {
            band_step<FMT, PAT, MODE, HC, ODD, EPI, true, FMT == FMT_U8 ? 2 : 1>(f, sp, x, yA, stage, Pw, Dw,
                                                                              col_ok, col_core, lut);
        }
    }
}

}  // namespace

// Group -> (strip, group row).  Blocks are spread round-robin over the 8
// XCDs; the remap (xcd_remap bit 0) gives each XCD a contiguous run of
// groups, so the raw rows that vertically adjacent groups share are read
// through one L2.  XCDs 4-7 take their runs in reverse, so the frame's last
// rows, whose groups run the slower per-wave rim path, start first instead
// of finishing the launch (12 MP MG frame: 16.1 -> 15.8 us).  Strip-fastest
// (bit 1): consecutive groups cover a group row left to right; else they
// walk down a strip (place_split, with the host's place_mul).
__device__ __forceinline__ void group_place(const FastParams& f, int& sx, int& g) {
    const int nblk = gridDim.x;
    // with nblk = 8q + r, XCD x runs cnt = q + (x < r) blocks: a bijection of [0, nblk)
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int q8 = nblk >> 3, r8 = nblk & 7;
    const int cnt = q8 + (xcd < r8 ? 1 : 0);
    const int run = xcd >= 4 ? cnt - 1 - slot : slot;
    const uint32_t lblk = (uint32_t)__builtin_amdgcn_readfirstlane(
        (f.xcd_remap & 1) ? xcd * q8 + imin(xcd, r8) + run : (int)blockIdx.x);
    place_split(lblk, (f.xcd_remap & 2) != 0, (uint32_t)f.nsx, (uint32_t)f.nby, f.place_mul, sx, g);
}

// Wave w (wave-uniform) of a group on the shared-LDS path.
template <int FMT, int PAT, int MODE, int TH, int GB, bool ODD, bool EPI, bool LENS, bool HC>
__device__ __forceinline__ void group_waves(int w, const FastParams& f, const Strip& sp, int yg, uint8_t* L,
                                            const uint16_t* lut) {
    switch (w) {
    case 0: group_wave<FMT, PAT, MODE, TH, GB, ODD, EPI, LENS, 0, HC>(f, sp, yg, L, lut); break;
    case 1: group_wave<FMT, PAT, MODE, TH, GB, ODD, EPI, LENS, 1, HC>(f, sp, yg, L, lut); break;
    case 2: if (GB > 2) group_wave<FMT, PAT, MODE, TH, GB, ODD, EPI, LENS, (GB > 2 ? 2 : 0), HC>(f, sp, yg, L, lut); break;
    default: if (GB > 3) group_wave<FMT, PAT, MODE, TH, GB, ODD, EPI, LENS, (GB > 3 ? 3 : 0), HC>(f, sp, yg, L, lut); break;
    }
}

// One workgroup item (of the single-frame kernel or of the batched one): the
// group row from yg of strip sp, as wave w (wave-uniform).  Vertical rim:
// the group's outputs touch the frame's top or bottom 3-pixel rim, or its
// raw window reaches the stream's final row; its waves take their bands one
// by one on the per-wave path (vrim: the band's rows reach the rim, else
// only the strip's columns do).  Groups whose strip reaches the left or
// right rim run the shared path in its careful-column form.  L holds
// GroupLayout<TH, GB>::kBytes.  No barrier is left pending on return.
template <int FMT, int PAT, int MODE, int TH, int GB, bool ODD, bool EPI, bool LENS>
__device__ __forceinline__ void group_item(const FastParams& f, const Strip& sp, int yg, int w, uint8_t* L,
                                           const uint16_t* lut) {
    using G = GroupLayout<TH, GB>;
    const bool hrim = sp.xs < 3 || sp.xs + 248 > f.w - 3;
    if (yg < 3 || yg + G::kRows + 3 > f.h - 1) {
        const int y0 = yg + TH * w;
        if (y0 >= f.h) return;
        uint32_t* stage = (uint32_t*)L + w * 384;  // 1536 bytes of the group's LDS per wave
        const bool vrim = y0 < 3 || y0 + TH + 3 > f.h - 1;
        if (vrim) run_band<FMT, PAT, MODE, true, TH, ODD, EPI, true, LENS>(f, sp, y0, stage, lut);
        else if (hrim) run_band<FMT, PAT, MODE, true, TH, ODD, EPI, false, LENS>(f, sp, y0, stage, lut);
        else run_band<FMT, PAT, MODE, false, TH, ODD, EPI, false, LENS>(f, sp, y0, stage, lut);
        return;
    }
    if (hrim) group_waves<FMT, PAT, MODE, TH, GB, ODD, EPI, LENS, true>(w, f, sp, yg, L, lut);
    else group_waves<FMT, PAT, MODE, TH, GB, ODD, EPI, LENS, false>(w, f, sp, yg, L, lut);
}

template <int FMT, int PAT, int MODE, int TH, int GB, bool ODD, bool EPI = false, bool LENS = false>
__global__ __launch_bounds__(64 * GB) void k_fast(FastParams f) {
    using G = GroupLayout<TH, GB>;
    __shared__ uint4 s_group[G::kBytes / 16];
    // epilogue form: the part of the gamma table the demosaic's values can
    // reach (lut_last + 1 <= 4096 entries, checked on the host) in LDS,
    // filled by the whole workgroup before any wave leaves
    __shared__ __align__(16) uint16_t s_lut[EPI ? 4096 : 8];
    if (EPI && (f.epi.flags & EPI_GAMMA)) {
        // in 16-byte pieces: entries 0 .. lut_last rounded up to a multiple of
        // 8, inside the table (256 or 65536 entries, 256-byte aligned)
        const int n16 = ((int)lut_last(f) + 8) >> 3;
        const uint4* __restrict__ src = (const uint4*)f.epi.lut;
        for (int i = (int)threadIdx.x; i < n16; i += (int)blockDim.x) ((uint4*)s_lut)[i] = src[i];
        __syncthreads();
    }
    int sx, g;
    group_place(f, sx, g);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    group_item<FMT, PAT, MODE, TH, GB, ODD, EPI, LENS>(f, strip_of(f, sx), g * G::kRows, w, (uint8_t*)s_group,
                                                        s_lut);
}

// ---------------------------------------------------------------- launchers

template <int FMT, int PAT, int MODE, bool ODD>
static void launch_fmpo(const FastParams& f, hipStream_t st) {
    // f.nsx strips x f.nby groups of 4 waves x f.th rows
    const int blocks = f.nsx * f.nby;
    // the ISP forms (RGB epilogue, with or without the lens prologue): 4-row waves in groups of 4
    if (f.lens.on) {
        hipLaunchKernelGGL((k_fast<FMT, PAT, MODE, 4, 4, ODD, true, true>), dim3(blocks), dim3(256), 0, st, f);
        return;
    }
    if (f.epi.flags) {
        hipLaunchKernelGGL((k_fast<FMT, PAT, MODE, 4, 4, ODD, true>), dim3(blocks), dim3(256), 0, st, f);
        return;
    }
    if (f.th == 2) hipLaunchKernelGGL((k_fast<FMT, PAT, MODE, 2, 4, ODD>), dim3(blocks), dim3(256), 0, st, f);
    else hipLaunchKernelGGL((k_fast<FMT, PAT, MODE, 4, 4, ODD>), dim3(blocks), dim3(256), 0, st, f);
}

template <int FMT, int PAT, int MODE>
static hipError_t launch_fmp(const FastParams& f, hipStream_t st) {
    if (f.w % 4 == 2) launch_fmpo<FMT, PAT, MODE, true>(f, st);
    else launch_fmpo<FMT, PAT, MODE, false>(f, st);
    return hipGetLastError();
}

// The reference's 8-bit integer MG runs on 8-bit samples demosaiced from an
// unpacked buffer only, so only the u8 layout instantiates it.
template <int FMT, int PAT>
static hipError_t launch_fp(const FastParams& f, int mode, hipStream_t st) {
    switch (mode) {
    case MODE_HQLI: return launch_fmp<FMT, PAT, MODE_HQLI>(f, st);
    case MODE_MG_EXACT: return launch_fmp<FMT, PAT, MODE_MG_EXACT>(f, st);
    default:
        if constexpr (FMT == FMT_U8) return launch_fmp<FMT, PAT, MODE_MG_INT8>(f, st);
        else return hipErrorInvalidValue;
    }
}

template <int FMT>
static hipError_t launch_f(const FastParams& f, int mode, hipStream_t st) {
    switch (f.pattern) {
    case 0: return launch_fp<FMT, 0>(f, mode, st);
    case 1: return launch_fp<FMT, 1>(f, mode, st);
    case 2: return launch_fp<FMT, 2>(f, mode, st);
    default: return launch_fp<FMT, 3>(f, mode, st);
    }
}

}  // namespace imop_hip
