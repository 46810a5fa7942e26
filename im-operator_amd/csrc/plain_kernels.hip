// plain_kernels.hip -- fused unpack -> black level -> HQLI / MG demosaic for
// gfx950, one launch per frame, interleaved RGB written once.
//
// Tiling: a workgroup owns a TILE_W x TILE_H block of output pixels.  Phase 1
// decodes the raw samples of the block plus an apron into LDS (black level
// folded in).  For MG, phase 2 computes one "colour difference" per R/B site
// of the block +-1 into a second LDS plane.  Phase 3 gives every thread a
// column of 2x2 CFA quads and writes their RGB.
//
// Exactness: the reference MG >8-bit path (mg.cpp:163-216, avx2.cpp:606-686)
// runs in float, but every value it forms is a multiple of 1/32 far below
// 2^24, so it is reproduced here in integers:
//   G8  = 8 * G          (raw at G sites, 4c + 2*near - far at interior R/B,
//                         -8 for the -1 sentinel at border R/B sites)
//   D   = 8 * raw - G8   at R/B sites
//   R/B at a G site  = (16 * raw + D_a + D_b + 8) >> 4      (two same-colour 4-neighbours)
//   opposite colour at R/B = (32 * raw - 4 * D + sum of 4 diagonal D + 16) >> 5
//   G at R/B         = (G8 + 4) >> 3
// which equals clamp(int(v + 0.5f)) of the float result (write_float_planes
// _to_rgb, common.hpp:201-214) for every input.  The 8-bit integer path keeps
// the reference's own integer rounding, including C's truncating division.
//
// Pixels in the outer two rows / columns follow fill_intermediate_borders /
// fill_rgb_borders (common.hpp:47-129): they take the value the planes hold at
// a remapped position before the fill.  Those few pixels are computed by a
// per-pixel path that reads HBM directly.
#include "device_common.hpp"
#include "plain.hpp"

namespace imop_hip {

// ---------------------------------------------------------------- border path

// Source position of the border fill with width b (fill_intermediate_borders_t,
// common.hpp:98-129; fill_rgb_borders :47-96).
__device__ __forceinline__ void border_src(int x, int y, int w, int h, int b, int& sx, int& sy) {
    if (y < b) { sy = b; sx = x; }
    else if (y >= h - b) { sy = h - 1 - b; sx = x; }
    else { sy = y; sx = x < b ? b : (x >= w - b ? w - 1 - b : x); }
}

__device__ __forceinline__ int tdiv4(int v) { return (v + ((v >> 31) & 3)) >> 2; }   // C's v / 4
__device__ __forceinline__ int tdiv2(int v) { return (v + (int)((unsigned)v >> 31)) >> 1; }  // C's v / 2

__device__ __forceinline__ bool g_interior(int x, int y, int w, int h) {
    return x >= 2 && x < w - 2 && y >= 2 && y < h - 2;
}

__device__ __forceinline__ int gnum_at(const FrameParams& f, int x, int y) {
    int c = sample_at(f, x, y);
    int nr = sample_at(f, x, y - 1) + sample_at(f, x, y + 1) + sample_at(f, x + 1, y) + sample_at(f, x - 1, y);
    int fr = sample_at(f, x, y - 2) + sample_at(f, x, y + 2) + sample_at(f, x + 2, y) + sample_at(f, x - 2, y);
    return 4 * c + 2 * nr - fr;
}

// MG exact: 8 * G plane value before the border fill.
__device__ int mg_g8(const FrameParams& f, int pat, int x, int y) {
    if (site_of(pat, y, x) == 1) return 8 * sample_at(f, x, y);
    if (g_interior(x, y, f.w, f.h)) return gnum_at(f, x, y);
    return -8;
}

// MG 8-bit: G plane value before the border fill (AVX2 variant: sentinel 0).
__device__ int mg_gi(const FrameParams& f, int pat, int x, int y) {
    if (site_of(pat, y, x) == 1) return sample_at(f, x, y);
    if (g_interior(x, y, f.w, f.h)) return iclamp((gnum_at(f, x, y) + 4) >> 3, 0, 255);
    return 0;
}

// MG output pixel at a position (x, y) of the pre-fill planes, rounded.
__device__ void mg_pre_rgb(const FrameParams& f, int pat, int mode, int x, int y, int& r, int& g, int& b) {
    const int w = f.w, h = f.h;
    const int s = site_of(pat, y, x);
    const bool core = x >= 1 && x < w - 1 && y >= 1 && y < h - 1;
    if (mode == MODE_MG_EXACT) {
        int g8 = mg_g8(f, pat, x, y);
        int r32 = 0, b32 = 0;
        if (core) {
            int p = sample_at(f, x, y);
            if (s != 1) {
                int sum = 0;
                for (int k = 0; k < 4; ++k) {
                    int xx = x + ((k & 1) ? 1 : -1), yy = y + ((k & 2) ? 1 : -1);
                    sum += 8 * sample_at(f, xx, yy) - mg_g8(f, pat, xx, yy);
                }
                int other = 4 * g8 + sum;
                if (s == 0) { r32 = 32 * p; b32 = other; } else { b32 = 32 * p; r32 = other; }
            } else {
                int rs = 0, bs = 0;
                for (int k = 0; k < 4; ++k) {
                    int xx = x + (k == 2 ? -1 : k == 3 ? 1 : 0), yy = y + (k == 0 ? -1 : k == 1 ? 1 : 0);
                    int t = site_of(pat, yy, xx);
                    int d = 8 * sample_at(f, xx, yy) - mg_g8(f, pat, xx, yy);
                    if (t == 0) rs += d; else if (t == 2) bs += d;
                }
                r32 = 4 * g8 + 2 * rs;
                b32 = 4 * g8 + 2 * bs;
            }
        }
        r = iclamp((r32 + 16) >> 5, 0, f.clamp_hi);
        g = iclamp((g8 + 4) >> 3, 0, f.clamp_hi);
        b = iclamp((b32 + 16) >> 5, 0, f.clamp_hi);
    } else {
        int gi = mg_gi(f, pat, x, y);
        int rr = 0, bb = 0;
        if (core) {
            int p = sample_at(f, x, y);
            if (s != 1) {
                int sum = 0;
                for (int k = 0; k < 4; ++k) {
                    int xx = x + ((k & 1) ? 1 : -1), yy = y + ((k & 2) ? 1 : -1);
                    sum += sample_at(f, xx, yy) - mg_gi(f, pat, xx, yy);
                }
                int other = gi + tdiv4(sum + 2);
                if (s == 0) { rr = p; bb = other; } else { bb = p; rr = other; }
            } else {
                int rs = 0, bs = 0;
                for (int k = 0; k < 4; ++k) {
                    int xx = x + (k == 2 ? -1 : k == 3 ? 1 : 0), yy = y + (k == 0 ? -1 : k == 1 ? 1 : 0);
                    int t = site_of(pat, yy, xx);
                    int d = sample_at(f, xx, yy) - mg_gi(f, pat, xx, yy);
                    if (t == 0) rs += d; else if (t == 2) bs += d;
                }
                rr = gi + tdiv2(rs + 1);
                bb = gi + tdiv2(bs + 1);
            }
        }
        r = iclamp(rr, 0, 255);
        g = iclamp(gi, 0, 255);
        b = iclamp(bb, 0, 255);
    }
}

// HQLI interior pixel (hqli_core, interp_core.hpp:28-68).
__device__ void hqli_px(const FrameParams& f, int pat, int x, int y, int& r, int& g, int& b) {
    const int s = site_of(pat, y, x);
    const int c = sample_at(f, x, y);
    if (s != 1) {
        int cross = sample_at(f, x - 1, y) + sample_at(f, x + 1, y) + sample_at(f, x, y - 1) + sample_at(f, x, y + 1);
        int diag = sample_at(f, x - 1, y - 1) + sample_at(f, x + 1, y - 1) + sample_at(f, x - 1, y + 1) +
                   sample_at(f, x + 1, y + 1);
        int own = c, gg = (cross + 2) >> 2, oth = (diag + 2) >> 2;
        if (s == 0) { r = own; g = gg; b = oth; } else { b = own; g = gg; r = oth; }
    } else {
        int hz = (sample_at(f, x - 1, y) + sample_at(f, x + 1, y) + 1) >> 1;
        int vt = (sample_at(f, x, y - 1) + sample_at(f, x, y + 1) + 1) >> 1;
        g = c;
        if ((y & 1) == r_row_of(pat)) { r = hz; b = vt; } else { r = vt; b = hz; }
    }
    r = imin(r, f.clamp_hi); g = imin(g, f.clamp_hi); b = imin(b, f.clamp_hi);
}

// Any output pixel, straight from HBM (used for the outer ring).
__device__ void border_pixel(const FrameParams& f, int pat, int mode, int x, int y) {
    int r, g, b, sx, sy;
    border_src(x, y, f.w, f.h, 2, sx, sy);
    if (mode == MODE_HQLI) {
        // Rows copied whole from row 2 / h-3 keep that row's own edge
        // pixels as they were before the column fill: never written by the
        // interior pass, so 0 in a zero-initialised buffer.
        if (sx < 2 || sx >= f.w - 2) { r = g = b = 0; }
        else hqli_px(f, pat, sx, sy, r, g, b);
    } else {
        mg_pre_rgb(f, pat, mode, sx, sy, r, g, b);
    }
    put_px(f, x, y, r, g, b);
}

// ---------------------------------------------------------------- tiled path

template <int A, int LW, int LH>
__device__ __forceinline__ void fill_tile(const FrameParams& f, int* __restrict__ sp, int x0, int y0) {
    for (int k = threadIdx.x; k < LW * LH; k += blockDim.x) {
        int ly = k / LW, lx = k - ly * LW;
        int gx = x0 - A + lx, gy = y0 - A + ly;
        int v = 0;
        if (gx >= 0 && gx < f.w && gy >= 0 && gy < f.h) v = sample_at(f, gx, gy);
        sp[k] = v;
    }
}

// Store one 2-pixel quad row.  Even x and even width give 4-byte (u16) or
// 2-byte (u8) aligned runs.
__device__ __forceinline__ void store_pair(const FrameParams& f, int x, int y, int r0, int g0, int b0, int r1,
                                           int g1, int b1) {
    const uint64_t i = ((uint64_t)y * (uint64_t)f.w + (uint64_t)x) * 3u;
    if (f.out16) {
        uint32_t* o = (uint32_t*)((uint16_t*)f.out + i);
        o[0] = (uint32_t)r0 | ((uint32_t)g0 << 16);
        o[1] = (uint32_t)b0 | ((uint32_t)r1 << 16);
        o[2] = (uint32_t)g1 | ((uint32_t)b1 << 16);
    } else {
        uint16_t* o = (uint16_t*)(f.out + i);
        o[0] = (uint16_t)(r0 | (g0 << 8));
        o[1] = (uint16_t)(b0 | (r1 << 8));
        o[2] = (uint16_t)(g1 | (b1 << 8));
    }
}

// One output pixel of a quad from the LDS tiles: sample plane `sp` (apron A)
// and, for MG, colour-difference plane `sd` (apron 1).  (a, bb) is the
// pixel's offset inside the quad, a compile-time constant after unrolling.
template <int PAT, int MODE, int A, int LW, int DW>
__device__ __forceinline__ void quad_px(const int* __restrict__ sp, const int* __restrict__ sd, int lq, int qx, int a,
                                        int bb, int clamp_hi, int& r, int& g, int& b) {
    const int s = site_of(PAT, a, bb);
    const int* c = sp + (2 * lq + a + A) * LW + (2 * qx + bb + A);
    if (MODE == MODE_HQLI) {
        if (s != 1) {
            int cross = c[-1] + c[1] + c[-LW] + c[LW];
            int diag = c[-LW - 1] + c[-LW + 1] + c[LW - 1] + c[LW + 1];
            int gg = (cross + 2) >> 2, oth = (diag + 2) >> 2;
            if (s == 0) { r = c[0]; g = gg; b = oth; } else { b = c[0]; g = gg; r = oth; }
        } else {
            int hz = (c[-1] + c[1] + 1) >> 1, vt = (c[-LW] + c[LW] + 1) >> 1;
            g = c[0];
            if (a == r_row_of(PAT)) { r = hz; b = vt; } else { r = vt; b = hz; }
        }
        r = imin(r, clamp_hi); g = imin(g, clamp_hi); b = imin(b, clamp_hi);
    } else {
        const int* d = sd + (2 * lq + a + 1) * DW + (2 * qx + bb + 1);
        const int p = c[0];
        if (s != 1) {
            const int dsum = d[-DW - 1] + d[-DW + 1] + d[DW - 1] + d[DW + 1];
            int gg, oth;
            if (MODE == MODE_MG_EXACT) {
                gg = (8 * p - d[0] + 4) >> 3;
                oth = (32 * p - 4 * d[0] + dsum + 16) >> 5;
            } else {
                gg = p - d[0];
                oth = gg + tdiv4(dsum + 2);
            }
            if (s == 0) { r = p; g = gg; b = oth; } else { b = p; g = gg; r = oth; }
        } else {
            int hs = d[-1] + d[1], vs = d[-DW] + d[DW];
            int hz, vt;
            if (MODE == MODE_MG_EXACT) {
                hz = (16 * p + hs + 8) >> 4;
                vt = (16 * p + vs + 8) >> 4;
            } else {
                hz = p + tdiv2(hs + 1);
                vt = p + tdiv2(vs + 1);
            }
            g = p;
            if (a == r_row_of(PAT)) { r = hz; b = vt; } else { r = vt; b = hz; }
        }
        r = iclamp(r, 0, clamp_hi); g = iclamp(g, 0, clamp_hi); b = iclamp(b, 0, clamp_hi);
    }
}

template <int PAT, int MODE, int QPT>
__global__ __launch_bounds__(256) void k_plain_tile(FrameParams f, int tiles_x) {
    constexpr int TW = 128, TH = 8 * QPT;
    constexpr int A = (MODE == MODE_HQLI) ? 1 : 3;
    constexpr int LW = TW + 2 * A, LH = TH + 2 * A;
    constexpr int DW = TW + 2, DH = TH + 2;
    constexpr int DSZ = (MODE == MODE_HQLI) ? 1 : DW * DH;
    __shared__ int sp[LW * LH];
    __shared__ int sd[DSZ];

    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int x0 = tx * TW, y0 = ty * TH;
    const int w = f.w, h = f.h;

    fill_tile<A, LW, LH>(f, sp, x0, y0);
    __syncthreads();

    if constexpr (MODE != MODE_HQLI) {
        for (int k = threadIdx.x; k < DW * DH; k += blockDim.x) {
            int dy = k / DW, dx = k - dy * DW;
            int gx = x0 - 1 + dx, gy = y0 - 1 + dy;
            if (site_of(PAT, gy, gx) == 1) continue;
            const int* c = sp + (dy + A - 1) * LW + (dx + A - 1);
            int v;
            if (g_interior(gx, gy, w, h)) {
                int gn = 4 * c[0] + 2 * (c[-LW] + c[LW] + c[1] + c[-1]) - (c[-2 * LW] + c[2 * LW] + c[2] + c[-2]);
                if (MODE == MODE_MG_EXACT) v = 8 * c[0] - gn;
                else v = c[0] - iclamp((gn + 4) >> 3, 0, 255);
            } else {
                v = (MODE == MODE_MG_EXACT) ? 8 * c[0] + 8 : c[0];
            }
            sd[k] = v;
        }
        __syncthreads();
    }

    const int qx = threadIdx.x & 63;
    const int qrow = threadIdx.x >> 6;
    const bool pair_ok = (w & 1) == 0 && (((uintptr_t)f.out & 3) == 0);
    for (int j = 0; j < QPT; ++j) {
        const int lq = j * 4 + qrow;  // quad row within the tile
        const int px = x0 + 2 * qx, py = y0 + 2 * lq;
        if (px >= w || py >= h) continue;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            int r0, g0, b0, r1, g1, b1;
            quad_px<PAT, MODE, A, LW, DW>(sp, sd, lq, qx, a, 0, f.clamp_hi, r0, g0, b0);
            quad_px<PAT, MODE, A, LW, DW>(sp, sd, lq, qx, a, 1, f.clamp_hi, r1, g1, b1);
            const int y = py + a;
            if (y < h) {
                // The outer two rows / columns are written by k_plain_border.
                const bool in0 = g_interior(px, y, w, h), in1 = g_interior(px + 1, y, w, h);
                if (in0 && in1 && pair_ok) {
                    store_pair(f, px, y, r0, g0, b0, r1, g1, b1);
                } else {
                    if (in0) put_px(f, px, y, r0, g0, b0);
                    if (in1) put_px(f, px + 1, y, r1, g1, b1);
                }
            }
        }
    }
}

// The outer ring (2 rows / columns) of a frame, one thread per pixel, read
// straight from HBM.  Ring pixel k: the top two rows, the bottom two rows,
// then the left and right two columns of the rows in between.
template <int PAT, int MODE>
__global__ __launch_bounds__(256) void k_plain_border(FrameParams f) {
    const int w = f.w, h = f.h;
    const long long ring = 4LL * w + 4LL * (h - 4);
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ring) return;
    int x, y;
    if (k < 4LL * w) {
        const int r = (int)(k / w);
        x = (int)(k - (long long)r * w);
        y = r < 2 ? r : h - 4 + r;
    } else {
        const long long m = k - 4LL * w;
        const int r = (int)(m >> 2), c = (int)(m & 3);
        y = 2 + r;
        x = c < 2 ? c : w - 4 + c;
    }
    border_pixel(f, PAT, MODE, x, y);
}

// ---------------------------------------------------------------- launchers

template <int PAT, int MODE>
static hipError_t launch_pm(const FrameParams& f, hipStream_t st) {
    constexpr int QPT = 4;
    const int tiles_x = (f.w + 127) / 128;
    const int tiles_y = (f.h + 8 * QPT - 1) / (8 * QPT);
    hipLaunchKernelGGL((k_plain_tile<PAT, MODE, QPT>), dim3(tiles_x * tiles_y), dim3(256), 0, st, f, tiles_x);
    const long long ring = 4LL * f.w + 4LL * (f.h - 4);
    hipLaunchKernelGGL((k_plain_border<PAT, MODE>), dim3((unsigned)((ring + 255) / 256)), dim3(256), 0, st, f);
    return hipGetLastError();
}

template <int PAT>
static hipError_t launch_p(const FrameParams& f, int mode, hipStream_t st) {
    switch (mode) {
    case MODE_HQLI: return launch_pm<PAT, MODE_HQLI>(f, st);
    case MODE_MG_EXACT: return launch_pm<PAT, MODE_MG_EXACT>(f, st);
    default: return launch_pm<PAT, MODE_MG_INT8>(f, st);
    }
}

hipError_t launch_plain(const FrameParams& f, int mode, hipStream_t st) {
    switch (f.pattern) {
    case 0: return launch_p<0>(f, mode, st);
    case 1: return launch_p<1>(f, mode, st);
    case 2: return launch_p<2>(f, mode, st);
    default: return launch_p<3>(f, mode, st);
    }
}

}  // namespace imop_hip

namespace imop_hip {

// ---------------------------------------------------------------- black level

// In-place black level on an unpacked u8 / u16 frame (global.cpp:6-24,
// per_channel.cpp:6-33): raw value, minus the site's offset in float,
// truncated, clamped to [0, max].
__global__ __launch_bounds__(256) void k_black_level(uint8_t* data, int w, int h, int bd, int maxv, float o00,
                                                     float o01, float o10, float o11) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    const uint64_t i = (uint64_t)y * (uint64_t)w + (uint64_t)x;
    const float off = (y & 1) ? ((x & 1) ? o11 : o10) : ((x & 1) ? o01 : o00);
    if (bd <= 8) {
        int v = data[i];
        data[i] = (uint8_t)iclamp((int)((float)v - off), 0, maxv);
    } else {
        uint16_t* d = (uint16_t*)data;
        int v = d[i];
        d[i] = (uint16_t)iclamp((int)((float)v - off), 0, maxv);
    }
}

hipError_t launch_black_level(uint8_t* data, int w, int h, int bd, const float q[4], hipStream_t st) {
    const int maxv = bd > 16 ? 65535 : (int)((1u << bd) - 1u);
    dim3 blk(64, 4), grd((w + 63) / 64, (h + 3) / 4);
    hipLaunchKernelGGL(k_black_level, grd, blk, 0, st, data, w, h, bd, maxv, q[0], q[1], q[2], q[3]);
    return hipGetLastError();
}

}  // namespace imop_hip
