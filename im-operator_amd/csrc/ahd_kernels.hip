// ahd_kernels.hip -- AHD demosaic (src/algorithms/ahd.cpp:43-187) on gfx950,
// one LDS-tiled pass.
//
// The reference forms horizontal and vertical bilinear estimates (samples
// clamped into the frame), their colour differences L = R - G and M = B - G,
// filters each difference plane with the 3x3 exchange network on
// [1, w-1) x [1, h-1) (the outer ring keeps its values), and per pixel keeps
// the direction whose filtered L and M have the smaller sum of 3x3 variances
// (neighbour coordinates clamped; mean = sum / 9; var = the fused
// fmaf(d, d, var) GCC forms from `var += d * d`, / 9); a strict < picks
// horizontal.  Output (int)(v + 0.5f) clamped to [0, max].
//
// Every output depends on samples at most 3 pixels away, so one workgroup
// computes a 32 x 16 tile from LDS: the samples over the tile + 3, the four
// unfiltered difference planes over the tile + 2, the filtered planes over
// the tile + 1, then the variances, the choice and the RGB of its pixels.
// Plane entries are indexed by frame position; reads of positions outside
// the frame use the clamped position, as the reference does.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "median9.hpp"
#include "window_decode.hpp"
#include "plain.hpp"

#pragma clang fp contract(off)

namespace imop_hip {
namespace {

constexpr int TW = 32, TH = 16;
constexpr int RW = TW + 6, RH = TH + 6;  // samples: x0-3 .. x0+TW+2
constexpr int DW = TW + 4, DH = TH + 4;  // unfiltered differences: x0-2 .. x0+TW+1
constexpr int FW = TW + 2, FH = TH + 2;  // filtered differences: x0-1 .. x0+TW

// Horizontal (dir 0) or vertical (dir 1) estimate at a frame position from
// the staged samples S(dx, dy) (clamped positions), site s (0 R, 1 G, 2 B).
template <class Sf>
__device__ __forceinline__ void ahd_est(Sf S, int dir, int s, bool r_beside, float& r, float& g, float& b) {
    if (s == 1) {
        g = S(0, 0);
        const float hz = (S(-1, 0) + S(1, 0)) * 0.5f, vt = (S(0, -1) + S(0, 1)) * 0.5f;
        const float along = dir == 0 ? hz : vt, across = dir == 0 ? vt : hz;
        r = r_beside ? along : across;
        b = r_beside ? across : along;
    } else {
        const float own = S(0, 0);
        g = dir == 0 ? (S(-1, 0) + S(1, 0)) * 0.5f : (S(0, -1) + S(0, 1)) * 0.5f;
        const float diag = dir == 0 ? (S(-1, -1) + S(1, -1) + S(-1, 1) + S(1, 1)) * 0.25f
                                    : (S(-1, -1) + S(-1, 1) + S(1, -1) + S(1, 1)) * 0.25f;
        r = s == 0 ? own : diag;
        b = s == 0 ? diag : own;
    }
}

// At a G site: is R the chroma beside the pixel along dir (ahd.cpp:83, :110)?
__device__ __forceinline__ bool r_beside(int pat, int dir, int x, int y) {
    return dir == 0 ? (site_of(pat, y, x - 1) == 0 || site_of(pat, y, x + 1) == 0)
                    : (site_of(pat, y - 1, x) == 0 || site_of(pat, y + 1, x) == 0);
}

__global__ __launch_bounds__(256) void k_ahd(FrameParams f) {
    __shared__ float s_raw[RH][RW];
    __shared__ float s_d[4][DH][DW];  // Lh, Mh, Lv, Mv
    __shared__ float s_f[4][FH][FW];
    const int w = f.w, h = f.h, pat = f.pattern;
    const int ntx = (w + TW - 1) / TW;
    const int x0 = (blockIdx.x % ntx) * TW, y0 = (blockIdx.x / ntx) * TH;

    // samples at clamped positions; a window inside the frame (all but the
    // edge tiles) needs no clamp and takes the per-layout rows
    if (!window_inside<RW, RH>(f, x0 - 3, y0 - 3, [&](int j, int i, int v) { s_raw[j][i] = (float)v; }))
        for (int k = threadIdx.x; k < RH * RW; k += blockDim.x) {
            const int j = k / RW, i = k - j * RW;
            const int x = min(max(x0 - 3 + i, 0), w - 1), y = min(max(y0 - 3 + j, 0), h - 1);
            s_raw[j][i] = (float)sample_at(f, x, y);
        }
    __syncthreads();
    // unfiltered differences at frame positions x0-2+i, y0-2+j (in-frame ones are read)
    for (int k = threadIdx.x; k < DH * DW; k += blockDim.x) {
        const int j = k / DW, i = k - j * DW;
        const int x = x0 - 2 + i, y = y0 - 2 + j;
        if (x < 0 || x >= w || y < 0 || y >= h) continue;
        auto S = [&](int dx, int dy) { return s_raw[j + 1 + dy][i + 1 + dx]; };
        const int st = site_of(pat, y, x);
#pragma unroll
        for (int dir = 0; dir < 2; ++dir) {
            float r, g, b;
            ahd_est(S, dir, st, st == 1 && r_beside(pat, dir, x, y), r, g, b);
            s_d[2 * dir][j][i] = r - g;
            s_d[2 * dir + 1][j][i] = b - g;
        }
    }
    __syncthreads();
    // filtered differences at x0-1+i, y0-1+j: the network inside [1, w-1) x [1, h-1), the value on the ring
    // Each task is two vertically adjacent positions of one column (rows j0,
    // j0 + 1): the four row triples their windows span are sorted once, the
    // middle two shared (median9_of_sorted_rows).  A triple that reaches
    // outside the frame holds unwritten entries and is used by no position.
    static_assert(FH % 2 == 0, "AHD filter: row pairs");
    for (int t = threadIdx.x; t < (FH / 2) * FW; t += blockDim.x) {
        const int jp = t / FW, i = t - jp * FW, j0 = 2 * jp;
        const int x = x0 - 1 + i;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            float tr[4][3];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) tr[q][dx + 1] = s_d[p][j0 + q][i + 1 + dx];
                sort3(tr[q][0], tr[q][1], tr[q][2]);
            }
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int j = j0 + r, y = y0 - 1 + j;
                if (x < 0 || x >= w || y < 0 || y >= h) continue;
                if (x >= 1 && x < w - 1 && y >= 1 && y < h - 1) {
                    float v[9];
#pragma unroll
                    for (int q = 0; q < 3; ++q)
#pragma unroll
                        for (int c = 0; c < 3; ++c) v[3 * q + c] = tr[r + q][c];
                    s_f[p][j][i] = median9_of_sorted_rows(v);
                } else {
                    s_f[p][j][i] = s_d[p][j + 1][i + 1];
                }
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TH * TW; k += blockDim.x) {
        const int j = k / TW, i = k - j * TW;
        const int x = x0 + i, y = y0 + j;
        if (x >= w || y >= h) continue;
        // neighbour columns / rows, clamped, as indices into the filtered planes
        int cx[3], cy[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            cx[d] = min(max(x + d - 1, 0), w - 1) - (x0 - 1);
            cy[d] = min(max(y + d - 1, 0), h - 1) - (y0 - 1);
        }
        float var[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            float mean = 0.0f;
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) mean += s_f[p][cy[dy]][cx[dx]];
            mean /= 9.0f;
            float v = 0.0f;
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) {
                    const float d = s_f[p][cy[dy]][cx[dx]] - mean;
                    v = fmaf(d, d, v);
                }
            var[p] = v / 9.0f;
        }
        const int dir = ((var[0] + var[1]) < (var[2] + var[3])) ? 0 : 1;
        auto S = [&](int dx, int dy) { return s_raw[j + 3 + dy][i + 3 + dx]; };
        const int st = site_of(pat, y, x);
        float r, g, b;
        ahd_est(S, dir, st, st == 1 && r_beside(pat, dir, x, y), r, g, b);
        const int R = min((int)(r + 0.5f), f.maxv), G = min((int)(g + 0.5f), f.maxv), B = min((int)(b + 0.5f), f.maxv);
        put_px(f, x, y, max(R, 0), max(G, 0), max(B, 0));
    }
}

}  // namespace

hipError_t launch_ahd(const FrameParams& f, hipStream_t st) {
    const uint64_t tiles = (uint64_t)((f.w + TW - 1) / TW) * (uint64_t)((f.h + TH - 1) / TH);
    if (tiles > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ahd, dim3((unsigned)tiles), dim3(256), 0, st, f);
    return hipGetLastError();
}

}  // namespace imop_hip
