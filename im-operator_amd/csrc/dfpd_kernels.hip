// dfpd_kernels.hip -- DFPD demosaic (src/algorithms/dfpd.cpp:146-263 and its
// 8-bit core, dfpd.cpp:7-144) on gfx950.
//
// Green on [2, w-2) x [2, h-2): the sample at G sites, else the horizontal or
// vertical average plus half the Laplacian correction along the direction of
// smaller gradient (the mean of both averages on a tie).  Then, inside
// [3, w-3) x [3, h-3), the chroma from the mean diagonal (R/B sites) or cross
// (G sites) colour differences against that green.  Frames demosaiced from an
// unpacked 8-bit buffer follow the integer core: (a + b + 1) >> 1 halves,
// (sum + cnt / 2) / cnt means with C division, clamp to [0, 255].  The others
// follow the float path, where every value is a multiple of 1/16 and so
// exact: (int)(v + 0.5f), clamp to [0, max].  Then fill_rgb_borders with
// width 5 over the interior written from 3 on (k_rgb_borders): the border
// rows' columns 3-4 and w-5..w-4 keep row 5 / h-6's computed values.
//
// k_dfpd: a workgroup computes 64 x 16 pixels from LDS: the samples over the
// tile + 3, green over the tile + 1, then the chroma of its pixels.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "window_decode.hpp"
#include "plain.hpp"

#pragma clang fp contract(off)

namespace imop_hip {
namespace {

constexpr int TW = 64, TH = 16;
constexpr int RW = TW + 6, RH = TH + 6;  // samples: x0-3 .. x0+TW+2
constexpr int GW = TW + 2, GH = TH + 2;  // green: x0-1 .. x0+TW

__global__ __launch_bounds__(256) void k_dfpd(FrameParams f) {
    __shared__ int s_raw[RH][RW];
    __shared__ float s_g[GH][GW];  // the 8-bit core's integer green is held exactly as float
    const int w = f.w, h = f.h, pat = f.pattern;
    const bool ip = f.byte_path != 0;
    const int ntx = (w + TW - 1) / TW;
    const int x0 = (blockIdx.x % ntx) * TW, y0 = (blockIdx.x / ntx) * TH;

    // samples at clamped positions; a window inside the frame needs no clamp
    if (!window_inside<RW, RH>(f, x0 - 3, y0 - 3, [&](int j, int i, int v) { s_raw[j][i] = v; }))
        for (int k = threadIdx.x; k < RH * RW; k += blockDim.x) {
            const int j = k / RW, i = k - j * RW;
            const int x = min(max(x0 - 3 + i, 0), w - 1), y = min(max(y0 - 3 + j, 0), h - 1);
            s_raw[j][i] = sample_at(f, x, y);
        }
    __syncthreads();
    for (int k = threadIdx.x; k < GH * GW; k += blockDim.x) {
        const int j = k / GW, i = k - j * GW;
        const int x = x0 - 1 + i, y = y0 - 1 + j;
        if (x < 2 || x >= w - 2 || y < 2 || y >= h - 2) continue;  // never read
        auto S = [&](int dx, int dy) { return s_raw[j + 2 + dy][i + 2 + dx]; };
        const int c = S(0, 0);
        float g;
        if (site_of(pat, y, x) == 1) {
            g = (float)c;
        } else {
            const int xm1 = S(-1, 0), xp1 = S(1, 0), xm2 = S(-2, 0), xp2 = S(2, 0);
            const int ym1 = S(0, -1), yp1 = S(0, 1), ym2 = S(0, -2), yp2 = S(0, 2);
            const int gh = abs(xm1 - xp1) + abs(2 * c - xm2 - xp2), gv = abs(ym1 - yp1) + abs(2 * c - ym2 - yp2);
            if (ip) {
                const int g_h = (xm1 + xp1 + 1) >> 1, g_v = (ym1 + yp1 + 1) >> 1;
                int gi;
                if (gh < gv) gi = g_h + ((c - ((xm2 + xp2 + 1) >> 1) + 1) >> 1);
                else if (gv < gh) gi = g_v + ((c - ((ym2 + yp2 + 1) >> 1) + 1) >> 1);
                else gi = (g_h + g_v + 1) >> 1;
                g = (float)gi;
            } else {
                const float g_h = (float)(xm1 + xp1) * 0.5f, g_v = (float)(ym1 + yp1) * 0.5f;
                if (gh < gv) g = g_h + ((float)c - (float)(xm2 + xp2) * 0.5f) * 0.5f;
                else if (gv < gh) g = g_v + ((float)c - (float)(ym2 + yp2) * 0.5f) * 0.5f;
                else g = (g_h + g_v) * 0.5f;
            }
        }
        s_g[j][i] = g;
    }
    __syncthreads();
    const int hi = ip ? 255 : f.maxv;
    for (int k = threadIdx.x; k < TH * TW; k += blockDim.x) {
        const int j = k / TW, i = k - j * TW;
        const int x = x0 + i, y = y0 + j;
        if (x < 3 || x >= w - 3 || y < 3 || y >= h - 3) continue;  // the border pass writes the rest
        auto S = [&](int dx, int dy) { return s_raw[j + 3 + dy][i + 3 + dx]; };
        auto G = [&](int dx, int dy) { return s_g[j + 1 + dy][i + 1 + dx]; };
        const int s = site_of(pat, y, x);
        const float g = G(0, 0);
        int R, Gc, B;
        if (ip) {
            const int gi = (int)g;
            if (s != 1) {
                // the four diagonal neighbours all hold the opposite colour
                const int sum = (S(-1, -1) - (int)G(-1, -1)) + (S(1, -1) - (int)G(1, -1)) +
                                (S(-1, 1) - (int)G(-1, 1)) + (S(1, 1) - (int)G(1, 1));
                const int other = gi + (sum + 2) / 4;
                R = s == 0 ? S(0, 0) : other;
                B = s == 0 ? other : S(0, 0);
            } else {
                // one chroma above and below, the other left and right
                const int vs = (S(0, -1) - (int)G(0, -1)) + (S(0, 1) - (int)G(0, 1));
                const int hs = (S(-1, 0) - (int)G(-1, 0)) + (S(1, 0) - (int)G(1, 0));
                const bool r_vert = site_of(pat, y - 1, x) == 0;
                R = gi + ((r_vert ? vs : hs) + 1) / 2;
                B = gi + ((r_vert ? hs : vs) + 1) / 2;
            }
            Gc = gi;
        } else {
            if (s != 1) {
                float sum = 0.0f;
                sum += (float)S(-1, -1) - G(-1, -1);
                sum += (float)S(1, -1) - G(1, -1);
                sum += (float)S(-1, 1) - G(-1, 1);
                sum += (float)S(1, 1) - G(1, 1);
                const int own = (int)((float)S(0, 0) + 0.5f);
                const int other = (int)(g + sum / 4.0f + 0.5f);
                R = s == 0 ? own : other;
                B = s == 0 ? other : own;
            } else {
                float vs = 0.0f, hs = 0.0f;
                vs += (float)S(0, -1) - G(0, -1);
                vs += (float)S(0, 1) - G(0, 1);
                hs += (float)S(-1, 0) - G(-1, 0);
                hs += (float)S(1, 0) - G(1, 0);
                const bool r_vert = site_of(pat, y - 1, x) == 0;
                R = (int)(g + (r_vert ? vs : hs) / 2.0f + 0.5f);
                B = (int)(g + (r_vert ? hs : vs) / 2.0f + 0.5f);
            }
            Gc = (int)(g + 0.5f);
        }
        put_px(f, x, y, min(max(R, 0), hi), min(max(Gc, 0), hi), min(max(B, 0), hi));
    }
}

}  // namespace

hipError_t launch_dfpd(const FrameParams& f, hipStream_t st) {
    const uint64_t tiles = (uint64_t)((f.w + TW - 1) / TW) * (uint64_t)((f.h + TH - 1) / TH);
    if (tiles > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dfpd, dim3((unsigned)tiles), dim3(256), 0, st, f);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    return launch_rgb_borders(f, 5, 3, st);
}

}  // namespace imop_hip
