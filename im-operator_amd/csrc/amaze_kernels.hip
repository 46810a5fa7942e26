// amaze_kernels.hip -- AMAZE demosaic (src/algorithms/amaze.cpp:158-200 and its
// 8-bit core, amaze.cpp:7-156) on gfx950, over three FP32 planes.
//
//   k_amaze_g    G plane: -1 everywhere outside [2, w-2) x [2, h-2); inside,
//                the sample at G sites, else the inverse-gradient blend of the
//                horizontal and vertical estimates (each with a quarter
//                Laplacian), fmaf(wh, gh, wv * gv) as GCC contracts
//                `wh * gh + wv * gv`.
//   k_amaze_rb   R and B inside [1, w-1) x [1, h-1): at R/B sites the sample
//                and the mean of the diagonal opposite-colour differences, at
//                G sites the means of the cross neighbours' differences
//                (common.hpp:143-199), each difference against the G plane
//                as it stands, the -1 outside the interior included; 0
//                elsewhere.
//   k_amaze_out  fill_intermediate_borders(2) by source remapping (rows 0-1
//                and h-2..h-1 copy row 2 / h-3 as it was before the column
//                fill), (int)(v + 0.5f) clamped to [0, max]; the 8-bit core
//                clamps to [0, 255] and then runs fill_rgb_borders(2) on its
//                output, which makes every border pixel the converted plane
//                value at its position clamped into [2, w-3) x [2, h-3).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "plain.hpp"

#pragma clang fp contract(off)

namespace imop_hip {
namespace {

struct AmazePlanes {
    float *g, *r, *b;
};

__device__ __forceinline__ AmazePlanes amaze_planes(void* ws, int w, int h) {
    const size_t n = ((size_t)w * (size_t)h + 63u) & ~(size_t)63u;
    float* p = (float*)ws;
    return AmazePlanes{p, p + n, p + 2 * n};
}

__device__ __forceinline__ float smp(const FrameParams& f, int x, int y) { return (float)sample_at(f, x, y); }

__global__ __launch_bounds__(256) void k_amaze_g(FrameParams f, void* ws) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int w = f.w, h = f.h;
    if (x >= w || y >= h) return;
    const AmazePlanes P = amaze_planes(ws, w, h);
    const size_t i = (size_t)y * (size_t)w + (size_t)x;
    float g = -1.0f;
    if (x >= 2 && x < w - 2 && y >= 2 && y < h - 2) {
        const float c = smp(f, x, y);
        if (site_of(f.pattern, y, x) == 1) {
            g = c;
        } else {
            const float hm1 = smp(f, x - 1, y), hp1 = smp(f, x + 1, y), hm2 = smp(f, x - 2, y), hp2 = smp(f, x + 2, y);
            const float vm1 = smp(f, x, y - 1), vp1 = smp(f, x, y + 1), vm2 = smp(f, x, y - 2), vp2 = smp(f, x, y + 2);
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
    P.g[i] = g;
}

__global__ __launch_bounds__(256) void k_amaze_rb(FrameParams f, void* ws) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int w = f.w, h = f.h;
    if (x >= w || y >= h) return;
    const AmazePlanes P = amaze_planes(ws, w, h);
    const size_t i = (size_t)y * (size_t)w + (size_t)x;
    float r = 0.0f, b = 0.0f;
    if (x >= 1 && x < w - 1 && y >= 1 && y < h - 1) {
        const float g = P.g[i];
        const int s = site_of(f.pattern, y, x);
        auto diff = [&](int nx, int ny) { return smp(f, nx, ny) - P.g[(size_t)ny * (size_t)w + (size_t)nx]; };
        if (s != 1) {
            float sum = 0.0f;
            int cnt = 0;
            const int dxs[4] = {-1, 1, -1, 1}, dys[4] = {-1, -1, 1, 1};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int nx = x + dxs[k], ny = y + dys[k];
                if (site_of(f.pattern, ny, nx) == 2 - s) {  // always inside the frame here
                    sum += diff(nx, ny);
                    ++cnt;
                }
            }
            const float other = g + (cnt > 0 ? sum / (float)cnt : 0.0f);
            const float own = smp(f, x, y);
            r = s == 0 ? own : other;
            b = s == 0 ? other : own;
        } else {
            float rs = 0.0f, bs = 0.0f;
            int rc = 0, bc = 0;
            const int dxs[4] = {0, 0, -1, 1}, dys[4] = {-1, 1, 0, 0};
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (site_of(f.pattern, y + dys[k], x + dxs[k]) == 0) {
                    rs += diff(x + dxs[k], y + dys[k]);
                    ++rc;
                }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (site_of(f.pattern, y + dys[k], x + dxs[k]) == 2) {
                    bs += diff(x + dxs[k], y + dys[k]);
                    ++bc;
                }
            r = g + (rc > 0 ? rs / (float)rc : 0.0f);
            b = g + (bc > 0 ? bs / (float)bc : 0.0f);
        }
    }
    P.r[i] = r;
    P.b[i] = b;
}

__device__ __forceinline__ int amaze_conv(float v, int hi) { return min(max((int)(v + 0.5f), 0), hi); }

__global__ __launch_bounds__(256) void k_amaze_out(FrameParams f, void* ws) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int w = f.w, h = f.h;
    if (x >= w || y >= h) return;
    const AmazePlanes P = amaze_planes(ws, w, h);
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
    const size_t s = (size_t)sy * (size_t)w + (size_t)sx;
    const int hi = f.byte_path ? 255 : f.maxv;
    put_px(f, x, y, amaze_conv(P.r[s], hi), amaze_conv(P.g[s], hi), amaze_conv(P.b[s], hi));
}

}  // namespace

size_t amaze_workspace_bytes(int w, int h) {
    return 3u * ((((size_t)w * (size_t)h + 63u) & ~(size_t)63u) * sizeof(float));
}

hipError_t launch_amaze(const FrameParams& f, void* ws, hipStream_t st) {
    const dim3 grid((unsigned)((f.w + 63) / 64), (unsigned)((f.h + 3) / 4)), block(256);
    if (grid.y > 65535u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_amaze_g, grid, block, 0, st, f, ws);
    hipLaunchKernelGGL(k_amaze_rb, grid, block, 0, st, f, ws);
    hipLaunchKernelGGL(k_amaze_out, grid, block, 0, st, f, ws);
    return hipGetLastError();
}

}  // namespace imop_hip
