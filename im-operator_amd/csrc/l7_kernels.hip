// l7_kernels.hip -- L7 demosaic (src/algorithms/l7.cpp:65-128; the AVX2 8-bit
// routine avx2.cpp:362-420 does the same arithmetic) on gfx950.
//
// Inside [3, w-3) x [3, h-3) each colour a pixel lacks is the rounded integer
// mean (sum + wsum / 2) / wsum of that colour's samples in the 7x7 window,
// weighted by the reference's table (centre 0); the pixel's own colour is its
// sample.  Frames demosaiced from an unpacked 8-bit buffer clamp to
// [0, 255], the others to [0, max].  Then fill_rgb_borders with width 3
// (common.hpp:47-96): top and bottom rows copy row 3 / h-4 between the side
// borders, the side borders copy the row's pixel 3 / w-4, and the four
// 3 x 3 corners are left at 0 (the reference copies output it never wrote
// there; a zero-filled buffer gives 0).
//
// k_l7: a workgroup of 8 rows x 256 pixels stages its samples (decoded,
// black level applied) with a 3-pixel ring in LDS.  Each thread computes 4
// consecutive pixels of a row from a 7 x 10 window read as 16-bit quads;
// with the row's parity and the CFA pattern fixed, the colour of every tap
// and the weight sums are compile-time, so the window reduces to
// multiply-adds and divisions by constants.  The RGB bytes go through LDS
// to 16-byte non-temporal stores.  k_rgb_borders then fills the border.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "window_decode.hpp"
#include "plain.hpp"

namespace imop_hip {
namespace {

constexpr int SW = 256, SH = 8;                  // pixels per workgroup
constexpr int LW = SW + 8, LH = SH + 6;          // staged samples: x0-4 .. x0+SW+3, y0-3 .. y0+SH+2

constexpr int kW[7][7] = {{1, 4, 9, 16, 9, 4, 1},   {4, 9, 16, 25, 16, 9, 4},   {9, 16, 25, 36, 25, 16, 9},
                          {16, 25, 36, 0, 36, 25, 16}, {9, 16, 25, 36, 25, 16, 9}, {4, 9, 16, 25, 16, 9, 4},
                          {1, 4, 9, 16, 9, 4, 1}};

// Colour (0 R, 1 G, 2 B) of the sample at row parity py, column parity px.
template <int PAT>
__device__ constexpr int colour(int py, int px) {
    return site_of(PAT, py & 1, px & 1);
}

// Sum of the weights of colour c around a pixel at parities (YP, XP).
template <int PAT, int YP, int XP>
__device__ constexpr int wsum(int c) {
    int s = 0;
    for (int dy = -3; dy <= 3; ++dy)
        for (int dx = -3; dx <= 3; ++dx)
            if (colour<PAT>(YP + dy, XP + dx) == c) s += kW[dy + 3][dx + 3];
    return s;
}

// Pixel K of the thread's four (x even), row parity YP: RGB from the window
// win[7][10] (columns x-3 .. x+6).
template <int PAT, int YP, int K>
__device__ __forceinline__ void l7_px(const uint32_t (&win)[7][10], int hi, int& R, int& G, int& B) {
    uint32_t s[3] = {0u, 0u, 0u};
#pragma unroll
    for (int dy = -3; dy <= 3; ++dy)
#pragma unroll
        for (int dx = -3; dx <= 3; ++dx) {
            const int wv = kW[dy + 3][dx + 3];
            if (wv == 0) continue;
            s[colour<PAT>(YP + dy, K + dx)] += win[dy + 3][K + dx + 3] * (uint32_t)wv;
        }
    constexpr int own = colour<PAT>(YP, K);
    constexpr uint32_t w0 = (uint32_t)wsum<PAT, YP, K & 1>(0), w1 = (uint32_t)wsum<PAT, YP, K & 1>(1),
                       w2 = (uint32_t)wsum<PAT, YP, K & 1>(2);
    const uint32_t c = win[3][K + 3];
    const uint32_t r = own == 0 ? c : (s[0] + w0 / 2u) / w0;
    const uint32_t g = own == 1 ? c : (s[1] + w1 / 2u) / w1;
    const uint32_t b = own == 2 ? c : (s[2] + w2 / 2u) / w2;
    R = min((int)r, hi);
    G = min((int)g, hi);
    B = min((int)b, hi);
}

template <int PAT, int YP>
__device__ __forceinline__ void l7_four(const uint32_t (&win)[7][10], int hi, int (&rgb)[4][3]) {
    l7_px<PAT, YP, 0>(win, hi, rgb[0][0], rgb[0][1], rgb[0][2]);
    l7_px<PAT, YP, 1>(win, hi, rgb[1][0], rgb[1][1], rgb[1][2]);
    l7_px<PAT, YP, 2>(win, hi, rgb[2][0], rgb[2][1], rgb[2][2]);
    l7_px<PAT, YP, 3>(win, hi, rgb[3][0], rgb[3][1], rgb[3][2]);
}

__global__ __launch_bounds__(256) void k_l7(FrameParams f) {
    __shared__ __attribute__((aligned(16))) uint16_t s[LH][LW];
    __shared__ __attribute__((aligned(16))) uint8_t o[SH][SW * 6];  // RGB bytes of each row segment
    const int w = f.w, h = f.h;
    const int x0 = blockIdx.x * SW, y0 = blockIdx.y * SH;
    // samples at clamped positions; a window inside the frame needs no clamp
    if (!window_inside<LW, LH>(f, x0 - 4, y0 - 3, [&](int j, int i, int v) { s[j][i] = (uint16_t)v; }))
        for (int k = threadIdx.x; k < LH * LW; k += blockDim.x) {
            const int j = k / LW, i = k - j * LW;
            const int x = min(max(x0 - 4 + i, 0), w - 1), y = min(max(y0 - 3 + j, 0), h - 1);
            s[j][i] = (uint16_t)sample_at(f, x, y);
        }
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int hi = f.byte_path ? 255 : f.maxv;
    const int es = f.out16 ? 6 : 3;
    const int xa = max(x0, 3), xb = min(x0 + SW, w - 3);  // columns this block writes
#pragma unroll
    for (int t = 0; t < SH / 4; ++t) {
        const int row = wave + 4 * t, y = y0 + row;
        const int x = x0 + 4 * lane;
        if (y >= 3 && y < h - 3 && x + 3 >= xa && x < xb) {
            uint32_t win[7][10];
#pragma unroll
            for (int dy = 0; dy < 7; ++dy) {
                // columns x-4 .. x+7 of row y-3+dy as three 16-bit quads
                const uint2* q = (const uint2*)&s[row + dy][4 * lane];
                uint32_t v[12];
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    const uint2 p = q[e];
                    v[4 * e] = p.x & 0xFFFFu;
                    v[4 * e + 1] = p.x >> 16;
                    v[4 * e + 2] = p.y & 0xFFFFu;
                    v[4 * e + 3] = p.y >> 16;
                }
#pragma unroll
                for (int e = 0; e < 10; ++e) win[dy][e] = v[e + 1];
            }
            int rgb[4][3];
            switch (f.pattern * 2 + (y & 1)) {
            case 0: l7_four<0, 0>(win, hi, rgb); break;
            case 1: l7_four<0, 1>(win, hi, rgb); break;
            case 2: l7_four<1, 0>(win, hi, rgb); break;
            case 3: l7_four<1, 1>(win, hi, rgb); break;
            case 4: l7_four<2, 0>(win, hi, rgb); break;
            case 5: l7_four<2, 1>(win, hi, rgb); break;
            case 6: l7_four<3, 0>(win, hi, rgb); break;
            default: l7_four<3, 1>(win, hi, rgb); break;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (f.out16) {
                    uint16_t* q = (uint16_t*)&o[row][(4 * lane + k) * 6];
                    q[0] = (uint16_t)rgb[k][0]; q[1] = (uint16_t)rgb[k][1]; q[2] = (uint16_t)rgb[k][2];
                } else {
                    uint8_t* q = &o[row][(4 * lane + k) * 3];
                    q[0] = (uint8_t)rgb[k][0]; q[1] = (uint8_t)rgb[k][1]; q[2] = (uint8_t)rgb[k][2];
                }
            }
        }
    }
    __syncthreads();
    if (xa >= xb) return;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int t = 0; t < SH / 4; ++t) {
        const int row = wave + 4 * t, y = y0 + row;
        if (y < 3 || y >= h - 3) continue;
        // bytes [b0, b1) of the row segment: columns xa .. xb-1
        const int b0 = (xa - x0) * es, b1 = (xb - x0) * es;
        uint8_t* base = f.out + ((uint64_t)y * (uint64_t)w + (uint64_t)x0) * (uint64_t)es;
        // head bytes up to a 16-byte boundary, whole chunks, tail bytes
        const int head = min(b1, b0 + (int)((16u - ((uintptr_t)(base + b0) & 15u)) & 15u));
        for (int b = b0 + lane; b < head; b += 64) base[b] = o[row][b];
        const int chunks = (b1 - head) >> 4;
        for (int c = lane; c < chunks; c += 64) {
            const int b = head + 16 * c;
            u32x4 v;
            if ((b & 15) == 0) v = *(const u32x4*)&o[row][b];
            else {
                uint8_t tmp[16];
#pragma unroll
                for (int e = 0; e < 16; ++e) tmp[e] = o[row][b + e];
                v = __builtin_bit_cast(u32x4, tmp);
            }
            __builtin_nontemporal_store(v, (u32x4*)(base + b));
        }
        for (int b = head + 16 * chunks + lane; b < b1; b += 64) base[b] = o[row][b];
    }
}

// fill_rgb_borders(rgb, w, h, bd, b) (common.hpp:47-96) on an output whose
// interior [m, w-m) x [m, h-m) (m <= b) is written and the rest never was
// (0 here).  The reference copies whole rows first -- rows above b and below
// h-1-b take row b / h-1-b as it stands, its never-written ends giving 0 --
// and then the side columns of rows [b, h-b); ROWS selects that step, and
// the two steps run as two launches in that order.
template <bool ROWS>
__global__ __launch_bounds__(256) void k_rgb_borders(FrameParams f, int b, int m) {
    const int w = f.w, h = f.h;
    const uint64_t top = (uint64_t)b * (uint64_t)w;     // per band of rows (top, bottom)
    const uint64_t side = (uint64_t)(h - 2 * b) * (uint64_t)b;  // per side
    const uint64_t n = ROWS ? 2u * top : 2u * side;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        int x, y, sx, sy;
        bool zero = false;
        if (ROWS) {
            const uint64_t k = i < top ? i : i - top;
            y = (int)(k / (uint64_t)w);
            x = (int)(k - (uint64_t)y * (uint64_t)w);
            if (i >= top) y += h - b;
            sy = i < top ? b : h - 1 - b;
            sx = x;
            zero = x < m || x >= w - m;
        } else {
            const uint64_t k = i;
            const uint64_t kk = k < side ? k : k - side;
            y = b + (int)(kk / (uint64_t)b);
            const int c = (int)(kk - (uint64_t)(y - b) * (uint64_t)b);
            x = k < side ? c : w - b + c;
            sx = k < side ? b : w - 1 - b;
            sy = y;
        }
        const uint64_t d = ((uint64_t)y * (uint64_t)w + (uint64_t)x) * 3u;
        const uint64_t src = ((uint64_t)sy * (uint64_t)w + (uint64_t)sx) * 3u;
        if (f.out16) {
            uint16_t* o = (uint16_t*)f.out;
            for (int e = 0; e < 3; ++e) o[d + e] = zero ? (uint16_t)0 : o[src + e];
        } else {
            for (int e = 0; e < 3; ++e) f.out[d + e] = zero ? (uint8_t)0 : f.out[src + e];
        }
    }
}

}  // namespace

hipError_t launch_l7(const FrameParams& f, hipStream_t st) {
    const dim3 grid((unsigned)((f.w + SW - 1) / SW), (unsigned)((f.h + SH - 1) / SH)), block(256);
    if (grid.y > 65535u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_l7, grid, block, 0, st, f);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    return launch_rgb_borders(f, 3, 3, st);
}

hipError_t launch_rgb_borders(const FrameParams& f, int b, int m, hipStream_t st) {
    if (b <= 0 || f.w <= 2 * b || f.h <= 2 * b) return hipSuccess;
    const uint64_t rows = 2u * (uint64_t)b * (uint64_t)f.w, sides = 2u * (uint64_t)b * (uint64_t)(f.h - 2 * b);
    const uint64_t br = (rows + 255u) / 256u, bs = (sides + 255u) / 256u;
    hipLaunchKernelGGL(k_rgb_borders<true>, dim3((unsigned)(br < 4096u ? br : 4096u)), dim3(256), 0, st, f, b, m);
    hipLaunchKernelGGL(k_rgb_borders<false>, dim3((unsigned)(bs < 4096u ? (bs ? bs : 1u) : 4096u)), dim3(256), 0, st,
                       f, b, m);
    return hipGetLastError();
}

}  // namespace imop_hip
