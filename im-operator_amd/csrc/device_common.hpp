// device_common.hpp -- frame parameters, raw-pixel decode and black level on
// gfx950.  Everything here works on one sample at a time; the tiled kernels
// use it for their LDS fill and the border paths use it directly against HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace imop_hip {

// Raw layouts the decoder understands (include/imop/pixel_utils.hpp:58-189).
enum Fmt : int {
    FMT_U8 = 0,    // one byte per pixel (bd <= 8, packed or not: get_raw reads bytes)
    FMT_U16 = 1,   // little-endian u16 per pixel, MSB-aligned samples shifted down
    FMT_P12 = 2,   // 2 px / 3 B
    FMT_P10 = 3,   // 4 px / 5 B
    FMT_PGEN = 4,  // LSB-first bitstream, any other depth
    FMT_G10 = 5,   // GigE Vision 10-bit packing: 2 px / 3 B
};

// Demosaic modes of the tiled kernels.
enum Mode : int {
    MODE_HQLI = 0,    // bilinear, integer averages (hqli.cpp)
    MODE_MG_EXACT = 1,// MG float path, reproduced exactly in units of 1/32 (mg.cpp:163-216)
    MODE_MG_INT8 = 2, // MG 8-bit integer path, AVX2 variant (avx2.cpp:422-595)
};

// Everything a kernel needs about one frame.  Plain data; lives in kernel
// arguments (plain launches) or in a device-side table (batched launches).
struct FrameParams {
    const uint8_t* in;    // raw Bayer bytes (device)
    uint8_t* out;         // interleaved RGB, u8 if bd <= 8 else u16 (device)
    uint64_t in_bytes;    // size of the packed stream / unpacked buffer
    int32_t w, h;
    int32_t bd;
    int32_t fmt;          // Fmt
    int32_t maxv;         // (1 << bd) - 1
    int32_t clamp_hi;     // final clamp of the integer paths
    int32_t bl_on;        // black level applied before demosaic
    int32_t pattern;      // BayerPattern, 0..3
    int32_t out16;        // RGB written as u16 (bd > 8, or on request for 8-bit frames)
    float bl_q[4];        // offset by (y & 1) * 2 + (x & 1), already max(0, .)
    int32_t byte_path;    // 8-bit samples demosaiced from an unpacked buffer (the reference's 8-bit routines)
};

// Strips of the register-resident kernel.  A wave computes 248 columns
// [xs, xs + 248), xs = stride * strip.  Plain strips: stride 248, each strip
// writes its own columns.  Aligned strips (u16 output of the single-frame
// kernel): stride 224, and in row y the strip boundaries sit at
// stride * k + s(y), s(y) in {0, 8, 16, 24} chosen so that each boundary's
// output byte is 64-byte aligned; every write window then lies inside its
// wave's computed columns, and no 64-byte sector of the output is written
// half by one wave and half by another (such sectors stream at ~4.1 TB/s
// against ~5 TB/s for whole sectors).  The batched kernel keeps plain strips:
// its per-row scalar arithmetic would not fit the scalar registers of a
// kernel that inlines every layout.
__host__ __device__ inline int strip_stride(bool aligned) { return aligned ? 224 : 248; }
__host__ __device__ inline int strip_count(int w, bool aligned) {
    const int s = strip_stride(aligned);
    return w <= 248 ? 1 : (w - 248 + s - 1) / s + 1;
}

// Reciprocal of the item placement's divisor d (the strip or group-row
// count): q = (i * m) >> 40 equals i / d for every item i < n when n * d <
// 2^40 (that is, n < m) and n * m fits 64 bits.  m = (2^40 + e) / d with
// 0 <= e < d, so i * m / 2^40 = i / d + i * e / (d * 2^40), and the added
// term is below 1 / d -- the least gap from i / d up to the next integer --
// since i * e < n * d < 2^40.  0 when a bound fails: the kernels divide.
__host__ __device__ inline uint64_t place_reciprocal(uint64_t n, uint64_t d) {
    if (d == 0) return 0;
    const uint64_t m = ((1ull << 40) + d - 1) / d;
    return n < m && n <= ~0ull / m ? m : 0;
}

// Item i of a frame's nsx x nby items as (strip sx, group row or band g):
// strip-fastest (i = g * nsx + sx) or down each strip (i = sx * nby + g).
// m = place_reciprocal(items, divisor) or 0.  Unsigned throughout: the
// signed quotient and remainder of a plain `/` and `%` made the batched
// item 1.15x slower in the same kernel.
__host__ __device__ inline void place_split(uint32_t i, bool strip_fastest, uint32_t nsx, uint32_t nby, uint64_t m,
                                        int& sx, int& g) {
    const uint32_t d = strip_fastest ? nsx : nby;
    const uint32_t q = m ? (uint32_t)(((uint64_t)i * m) >> 40) : i / d;
    const uint32_t r = i - q * d;
    g = (int)(strip_fastest ? q : r);
    sx = (int)(strip_fastest ? r : q);
}

// RGB stages that may follow the demosaic in the same launch (epilogue.hpp).
enum EpilogueFlag : int { EPI_WB = 1, EPI_CCM = 2, EPI_GAMMA = 4 };

struct Epilogue {
    int32_t flags;         // EpilogueFlag bits; 0 = none
    int32_t maxv;          // detail::safe_max_val(bit depth)
    float gain[3];         // white balance gains r, g, b
    float m[9];            // colour matrix, row-major, out = M * in
    uint32_t lut_max;      // last gamma table entry: 255 (u8 RGB) or 65535 (u16 RGB)
    int32_t bounded;       // parameters finite and small: no stage value reaches 2^31
    const uint16_t* lut;   // gamma table for every storable value 0 .. lut_max (device)
};

// Polynomial lens shading applied to the decoded, black-levelled samples
// (lens_shading::process_polynomial, src/lens_shading/polynomial.cpp:11-52):
// the reference's per-call constants cx, cy, max_r as the host computes them,
// and the coefficient set of each CFA parity (y & 1) * 2 + (x & 1).
struct LensParams {
    float pc[4][3];     // {a2, a4, a6} of the channel at parity (y & 1) * 2 + (x & 1)
    float cx, cy, max_r;
    float inv_max_r;    // 1 / max_r correctly rounded
    int32_t on;
    // 1: the register-resident kernel's sequences are exact for every sample
    // of the frame (lens_px in fast_impl.hpp): correctly rounded sqrt and
    // division without range scaling, and products v * gain + 0.5 bounded
    // inside (-2^31, 2^31) so the clamp is a median
    int32_t fused_ok;
};

// Parameters of the register-resident HQLI / MG kernel (fast_kernels.hip).
struct FastParams {
    const uint8_t* in;
    uint8_t* out;
    uint64_t in_bytes;
    int32_t w, h, bd, maxv, clamp_hi, pattern;
    int32_t out16;        // RGB written as u16
    int32_t bl_on;
    int32_t bl_i[4];      // integral black level by (y & 1) * 2 + (x & 1)
    int32_t th;           // rows per wave (even)
    int32_t nsx, nby;     // strips (strip_count) and bands (single-frame kernel: groups of wpb bands)
    int32_t nwaves;
    int32_t wpb;          // waves per workgroup
    uint64_t place_mul;   // single-frame kernel: ceil(2^40 / divisor) of the group placement
    int32_t xcd_remap;    // give each XCD a contiguous run of waves
    int32_t aligned;      // aligned strips for u16 output (strip_stride)
    Epilogue epi;         // RGB stages applied before the stores (k_fast<.., EPI = true>)
    LensParams lens;      // lens shading of the decoded samples (k_fast<.., EPI = true>)
};

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int iclamp(int v, int lo, int hi) { return imin(imax(v, lo), hi); }

// Packed 12-bit: pixel i of a 3-byte pair group.  get_packed_raw's bound
// (pixel_utils.hpp:69-80): a group touching the stream's last byte reads 0.
__device__ __forceinline__ int decode_p12(const uint8_t* __restrict__ d, uint64_t i, uint64_t nb) {
    uint64_t o = (i >> 1) * 3u;
    if (o + 2u >= nb) return 0;
    unsigned b1 = d[o + 1];
    if ((i & 1u) == 0) return (int)(((unsigned)d[o] << 4) | (b1 & 0xFu));
    return (int)(((unsigned)d[o + 2] << 4) | (b1 >> 4));
}

// Packed 10-bit (pixel_utils.hpp:82-98): four high bytes, then the tails.
__device__ __forceinline__ int decode_p10(const uint8_t* __restrict__ d, uint64_t i, uint64_t nb) {
    uint64_t o = (i >> 2) * 5u;
    if (o + 4u >= nb) return 0;
    unsigned s = (unsigned)(i & 3u);
    return (int)(((unsigned)d[o + s] << 2) | (((unsigned)d[o + 4] >> (2u * s)) & 3u));
}

// GigE Vision 10-bit packing (BayerXX10Packed, the layout of the reference's
// test_data/*_pBayerRG10_Packed.raw): groups of 3 bytes hold two pixels,
// [P0 bits 9:2][P1 bits 1:0 in bits 5:4 | P0 bits 1:0 in bits 1:0][P1 bits 9:2].
// A group that runs past the stream decodes as 0.
__device__ __forceinline__ int decode_g10(const uint8_t* __restrict__ d, uint64_t i, uint64_t nb) {
    uint64_t o = (i >> 1) * 3u;
    if (o + 2u >= nb) return 0;
    unsigned b1 = d[o + 1];
    if ((i & 1u) == 0) return (int)(((unsigned)d[o] << 2) | (b1 & 3u));
    return (int)(((unsigned)d[o + 2] << 2) | ((b1 >> 4) & 3u));
}

// LSB-first bitstream (pixel_utils.hpp:100-117).
__device__ __forceinline__ int decode_pgen(const uint8_t* __restrict__ d, uint64_t i, uint64_t nb, int bd) {
    uint64_t bit = i * (uint64_t)bd;
    uint64_t o = bit >> 3;
    int sh = (int)(bit & 7u);
    if (o >= nb) return 0;
    int n = (bd + sh + 7) >> 3;
    if ((uint64_t)n > nb - o) n = (int)(nb - o);
    uint32_t v = 0;
    for (int k = 0; k < n; ++k) v |= (uint32_t)d[o + k] << (8 * k);
    return (int)((v >> sh) & ((1u << bd) - 1u));
}

// One raw sample as the demosaic sees it: decoded, then either black level
// (unpack -> process_global / process_per_channel -> demosaic of the
// unpacked buffer; black_level/global.cpp:6-24) or the u16 MSB-alignment of
// pixel::get_raw (pixel_utils.hpp:129-150).
__device__ __forceinline__ int sample_at(const FrameParams& f, int x, int y) {
    const uint64_t i = (uint64_t)y * (uint64_t)f.w + (uint64_t)x;
    int v;
    switch (f.fmt) {
    case FMT_U8: v = f.in[i]; break;
    case FMT_U16: v = (int)((const uint16_t*)f.in)[i]; break;
    case FMT_P12: v = decode_p12(f.in, i, f.in_bytes); break;
    case FMT_P10: v = decode_p10(f.in, i, f.in_bytes); break;
    case FMT_G10: v = decode_g10(f.in, i, f.in_bytes); break;
    default: v = decode_pgen(f.in, i, f.in_bytes, f.bd); break;
    }
    if (f.bl_on) {
        int nv = (int)((float)v - f.bl_q[((y & 1) << 1) | (x & 1)]);
        return iclamp(nv, 0, f.maxv);
    }
    if (f.fmt == FMT_U16 && f.bd > 8 && f.bd < 16 && v > f.maxv) v >>= (16 - f.bd);
    return v;
}

// Colour role of a site for a pattern: 0 = R, 1 = G, 2 = B
// (PatternOffsets::from_pattern, include/imop/types.hpp:88-101).
__host__ __device__ constexpr int r_row_of(int pat) { return (pat == 1 || pat == 3) ? 1 : 0; }
__host__ __device__ constexpr int r_col_of(int pat) { return (pat == 1 || pat == 2) ? 1 : 0; }
__host__ __device__ constexpr int site_of(int pat, int y, int x) {
    return (((y & 1) == r_row_of(pat)) && ((x & 1) == r_col_of(pat))) ? 0
         : (((y & 1) != r_row_of(pat)) && ((x & 1) != r_col_of(pat))) ? 2 : 1;
}

// Write one RGB pixel (u8, or u16 when out16).
__device__ __forceinline__ void put_px(const FrameParams& f, int x, int y, int r, int g, int b) {
    const uint64_t i = ((uint64_t)y * (uint64_t)f.w + (uint64_t)x) * 3u;
    if (!f.out16) {
        f.out[i] = (uint8_t)r; f.out[i + 1] = (uint8_t)g; f.out[i + 2] = (uint8_t)b;
    } else {
        uint16_t* o = (uint16_t*)f.out;
        o[i] = (uint16_t)r; o[i + 1] = (uint16_t)g; o[i + 2] = (uint16_t)b;
    }
}

}  // namespace imop_hip
