// amaze_kernels.hip -- AMAZE demosaic (src/algorithms/amaze.cpp:158-200 and its
// 8-bit core, amaze.cpp:7-156) on gfx950, one LDS-tiled pass (k_amaze).
//
//   G:       -1 everywhere outside [2, w-2) x [2, h-2); inside, the sample at
//            G sites, else the inverse-gradient blend of the horizontal and
//            vertical estimates (each with a quarter Laplacian),
//            fmaf(wh, gh, wv * gv) as GCC contracts `wh * gh + wv * gv`.
//   R, B:    inside [1, w-1) x [1, h-1): at R/B sites the sample and the mean
//            of the diagonal opposite-colour differences, at G sites the
//            means of the cross neighbours' differences (common.hpp:143-199),
//            each difference against G as it stands, the -1 outside the
//            interior included; 0 elsewhere.
//   output:  fill_intermediate_borders(2) by source remapping (rows 0-1 and
//            h-2..h-1 copy row 2 / h-3 as it was before the column fill),
//            (int)(v + 0.5f) clamped to [0, max]; the 8-bit core clamps to
//            [0, 255] and then runs fill_rgb_borders(2) on its output, which
//            makes every border pixel the converted value at its position
//            clamped into [2, w-3) x [2, h-3).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "window_decode.hpp"
#include "plain.hpp"

#pragma clang fp contract(off)

namespace imop_hip {
namespace {

__device__ __forceinline__ int amaze_conv(float v, int hi) { return min(max((int)(v + 0.5f), 0), hi); }

// The three steps fused for one 32 x 16 tile from LDS: samples over the tile
// + 3, G over the tile + 1, R and B over the tile, then the output through
// the border remap.  The last tile of each row / column is moved inward so
// that every tile lies inside the frame; a border pixel's source (at most 2
// pixels inward) is then in the tile that writes the pixel, and overlapping
// tiles write identical bytes.
constexpr int TW = 32, TH = 16;
constexpr int RW = TW + 6, RH = TH + 6;  // samples: x0-3 .. x0+TW+2
constexpr int GW = TW + 2, GH = TH + 2;  // G: x0-1 .. x0+TW

__global__ __launch_bounds__(256) void k_amaze(FrameParams f) {
    __shared__ float s_raw[RH][RW];
    __shared__ float s_g[GH][GW];
    __shared__ float s_r[TH][TW], s_b[TH][TW];
    const int w = f.w, h = f.h, pat = f.pattern;
    const int ntx = (w + TW - 1) / TW;
    const int x0 = min((int)(blockIdx.x % ntx) * TW, max(w - TW, 0));
    const int y0 = min((int)(blockIdx.x / ntx) * TH, max(h - TH, 0));

    // samples at clamped positions; a window inside the frame needs no clamp
    if (!window_inside<RW, RH>(f, x0 - 3, y0 - 3, [&](int j, int i, int v) { s_raw[j][i] = (float)v; }))
        for (int k = threadIdx.x; k < RH * RW; k += blockDim.x) {
            const int j = k / RW, i = k - j * RW;
            const int x = min(max(x0 - 3 + i, 0), w - 1), y = min(max(y0 - 3 + j, 0), h - 1);
            s_raw[j][i] = (float)sample_at(f, x, y);
        }
    __syncthreads();
    for (int k = threadIdx.x; k < GH * GW; k += blockDim.x) {
        const int j = k / GW, i = k - j * GW;
        const int x = x0 - 1 + i, y = y0 - 1 + j;
        float g = -1.0f;
        if (x >= 2 && x < w - 2 && y >= 2 && y < h - 2) {
            auto S = [&](int dx, int dy) { return s_raw[j + 2 + dy][i + 2 + dx]; };
            const float c = S(0, 0);
            if (site_of(pat, y, x) == 1) {
                g = c;
            } else {
                const float hm1 = S(-1, 0), hp1 = S(1, 0), hm2 = S(-2, 0), hp2 = S(2, 0);
                const float vm1 = S(0, -1), vp1 = S(0, 1), vm2 = S(0, -2), vp2 = S(0, 2);
                const float gH = fabsf(hm1 - hp1) + fabsf(2.0f * c - hm2 - hp2);
                const float gV = fabsf(vm1 - vp1) + fabsf(2.0f * c - vm2 - vp2);
                float wh = 1.0f / (1.0f + gH), wv = 1.0f / (1.0f + gV);
                const float sw = wh + wv;
                wh /= sw;
                wv /= sw;
                float gh = (hm1 + hp1) * 0.5f, gv = (vm1 + vp1) * 0.5f;
                const float lH = 2.0f * c - hm2 - hp2, lV = 2.0f * c - vm2 - vp2;
                gh += lH * 0.25f;
                gv += lV * 0.25f;
                g = fmaf(wh, gh, wv * gv);
            }
        }
        s_g[j][i] = g;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TH * TW; k += blockDim.x) {
        const int j = k / TW, i = k - j * TW;
        const int x = x0 + i, y = y0 + j;
        float r = 0.0f, b = 0.0f;
        if (x >= 1 && x < w - 1 && y >= 1 && y < h - 1) {
            auto S = [&](int dx, int dy) { return s_raw[j + 3 + dy][i + 3 + dx]; };
            auto G = [&](int dx, int dy) { return s_g[j + 1 + dy][i + 1 + dx]; };
            const float g = G(0, 0);
            const int s = site_of(pat, y, x);
            if (s != 1) {
                // all four diagonals hold the opposite colour
                float sum = 0.0f;
                sum += S(-1, -1) - G(-1, -1);
                sum += S(1, -1) - G(1, -1);
                sum += S(-1, 1) - G(-1, 1);
                sum += S(1, 1) - G(1, 1);
                const float other = g + sum / 4.0f;
                r = s == 0 ? S(0, 0) : other;
                b = s == 0 ? other : S(0, 0);
            } else {
                // one chroma above and below, the other left and right (in the reference's order)
                const float dn = S(0, -1) - G(0, -1), ds = S(0, 1) - G(0, 1);
                const float dw = S(-1, 0) - G(-1, 0), de = S(1, 0) - G(1, 0);
                float vs = 0.0f, hs = 0.0f;
                vs += dn;
                vs += ds;
                hs += dw;
                hs += de;
                const bool r_vert = site_of(pat, y - 1, x) == 0;
                r = g + (r_vert ? vs : hs) / 2.0f;
                b = g + (r_vert ? hs : vs) / 2.0f;
            }
        }
        s_r[j][i] = r;
        s_b[j][i] = b;
    }
    __syncthreads();
    const int hi = f.byte_path ? 255 : f.maxv;
    for (int k = threadIdx.x; k < TH * TW; k += blockDim.x) {
        const int j = k / TW, i = k - j * TW;
        const int x = x0 + i, y = y0 + j;
        if (x >= w || y >= h) continue;
        int sx, sy;
        if (f.byte_path) {
            sx = min(max(x, 2), w - 3);
            sy = min(max(y, 2), h - 3);
        } else if (y < 2 || y >= h - 2) {
            sy = y < 2 ? 2 : h - 3;
            sx = x;
        } else {
            sy = y;
            sx = min(max(x, 2), w - 3);
        }
        const int si = sx - x0, sj = sy - y0;
        put_px(f, x, y, amaze_conv(s_r[sj][si], hi), amaze_conv(s_g[sj + 1][si + 1], hi),
               amaze_conv(s_b[sj][si], hi));
    }
}

}  // namespace

hipError_t launch_amaze(const FrameParams& f, hipStream_t st) {
    const uint64_t tiles = (uint64_t)((f.w + TW - 1) / TW) * (uint64_t)((f.h + TH - 1) / TH);
    if (tiles > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_amaze, dim3((unsigned)tiles), dim3(256), 0, st, f);
    return hipGetLastError();
}

}  // namespace imop_hip
