// super_fast_kernels.hip -- SUPER_FAST demosaic (src/algorithms/super_fast.cpp:
// 9-43) on gfx950, with the reference's output on an AVX2 host
// (process_super_fast_optimized, avx2.cpp:99-253):
//   R/B site: own value, green = (l + r + u + d + 2) >> 2, the other chroma
//             from the diagonal neighbour (x +- 1 toward the pair, y +- 1);
//   G site:   own value, the row's chroma = (l + r + 1) >> 1, the column's
//             chroma = (u + d + 1) >> 1;
// every neighbour clamped into the frame.  Frames demosaiced from an unpacked
// 8-bit buffer follow super_fast_avx2_8bit (avx2.cpp:99-207), which works in
// 32-pixel vectors: a row's first vector takes its "left" operand from
// offset 0 (pixels 1..31 of the row see themselves as their left
// neighbour), and a vector starting at w - 32 takes "right" from itself;
// its bytes are stored unclamped.  Other frames clamp to the bit depth's
// maximum (set_rgb_clamped).
//
// One workgroup: 4 rows x 256 pixels; the samples (decoded, black level
// applied) of the block plus a 1-pixel ring are staged in LDS; each thread
// computes 4 consecutive pixels of one row into an LDS copy of the row's RGB
// bytes, which its wave then writes as contiguous 16-byte non-temporal
// stores (byte stores where the row segment does not start on 16 bytes).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "window_decode.hpp"
#include "plain.hpp"

namespace imop_hip {
namespace {

constexpr int SW = 256, SH = 4;              // pixels per workgroup
constexpr int LW = SW + 2, LH = SH + 2;      // staged samples: x0-1 .. x0+SW, y0-1 .. y0+SH

__global__ __launch_bounds__(256) void k_super_fast(FrameParams f) {
    __shared__ uint16_t s[LH][LW];
    __shared__ __attribute__((aligned(16))) uint8_t o[SH][SW * 6];  // RGB bytes of each row segment
    const int w = f.w, h = f.h;
    const int x0 = blockIdx.x * SW, y0 = blockIdx.y * SH;
    static_assert(SH * 64 == 256, "one wave per row");
    // samples at clamped positions; a window inside the frame needs no clamp
    if (!window_inside<LW, LH>(f, x0 - 1, y0 - 1, [&](int j, int i, int v) { s[j][i] = (uint16_t)v; }))
        for (int k = threadIdx.x; k < LH * LW; k += blockDim.x) {
            const int j = k / LW, i = k - j * LW;
            const int x = min(max(x0 - 1 + i, 0), w - 1), y = min(max(y0 - 1 + j, 0), h - 1);
            s[j][i] = (uint16_t)sample_at(f, x, y);
        }
    __syncthreads();

    const int row = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int y = min(y0 + row, h - 1);  // rows past the frame compute a copy of the last and store nothing
    const int full = (w / 32) * 32;  // pixels the 8-bit routine covers with 32-pixel vectors
    const bool gr_row = (y & 1) == r_row_of(f.pattern);
    const int ny = min(max((y & 1) ? y - 1 : y + 1, 0), h - 1);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int x = x0 + 4 * lane + k;
        if (x >= w) break;
        const int i = x - x0 + 1, j = row + 1;
        const int c = s[j][i];
        // clamped neighbours: at the frame's edge the staged ring repeats the edge sample
        int l = s[j][i - 1], r = s[j][i + 1];
        const int u = s[j - 1][i], d = s[j + 1][i];
        if (f.byte_path && x < full) {
            if (x < 32) l = c;
            if ((x & ~31) == w - 32) r = c;
        }
        const int st = site_of(f.pattern, y, x);
        int R, G, B;
        if (st != 1) {
            const int nx = min(max((x & 1) ? x - 1 : x + 1, 0), w - 1);
            const int diag = s[ny - y0 + 1][nx - x0 + 1];
            G = (l + r + u + d + 2) >> 2;
            R = st == 0 ? c : diag;
            B = st == 0 ? diag : c;
        } else {
            G = c;
            const int hz = (l + r + 1) >> 1, vt = (u + d + 1) >> 1;
            R = gr_row ? hz : vt;
            B = gr_row ? vt : hz;
        }
        if (!f.byte_path) {
            R = min(R, f.maxv);
            G = min(G, f.maxv);
            B = min(B, f.maxv);
        }
        if (f.out16) {
            uint16_t* q = (uint16_t*)&o[row][(x - x0) * 6];
            q[0] = (uint16_t)R; q[1] = (uint16_t)G; q[2] = (uint16_t)B;
        } else {
            uint8_t* q = &o[row][(x - x0) * 3];
            q[0] = (uint8_t)R; q[1] = (uint8_t)G; q[2] = (uint8_t)B;
        }
    }
    __syncthreads();
    if (y0 + row >= h) return;
    const int es = f.out16 ? 6 : 3;
    const int nbytes = min(SW, w - x0) * es;
    uint8_t* dst = f.out + ((uint64_t)y * (uint64_t)w + (uint64_t)x0) * (uint64_t)es;
    int done = 0;
    if (((uintptr_t)dst & 15u) == 0) {
        const int chunks = nbytes >> 4;
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        for (int c = lane; c < chunks; c += 64)
            __builtin_nontemporal_store(*(const u32x4*)&o[row][16 * c], (u32x4*)(dst + 16 * c));
        done = chunks << 4;
    }
    for (int b = done + lane; b < nbytes; b += 64) dst[b] = o[row][b];
}

}  // namespace

hipError_t launch_super_fast(const FrameParams& f, hipStream_t st) {
    const dim3 grid((unsigned)((f.w + SW - 1) / SW), (unsigned)((f.h + SH - 1) / SH)), block(256);
    if (grid.y > 65535u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_super_fast, grid, block, 0, st, f);
    return hipGetLastError();
}

}  // namespace imop_hip
