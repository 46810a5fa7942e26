// isp_kernels.hip -- the RGB epilogue (white balance, colour matrix, gamma;
// epilogue.hpp) as its own in-place pass over an interleaved RGB image.  The
// demosaic paths that cannot apply it in registers (PRISM, the generic
// LDS-tiled kernels) run it after the demosaic; it is also the standalone
// imop_hip_rgb_epilogue.
#include "epilogue.hpp"

namespace imop_hip {

namespace {

// 48 bytes per chunk: 8 u16 pixels or 16 u8 pixels, three 16-byte words.
// Gamma table lookups of the standalone pass: entries 0 .. lds_last from
// LDS, larger values (any u16 value above the bit depth's maximum) from the
// whole table in global memory.
struct LutSplit {
    const uint16_t* lds;
    uint32_t lds_last;
    const uint16_t* all;
    uint32_t all_last;
    __device__ __forceinline__ int operator()(int v) const {
        const uint32_t u = (uint32_t)v;
        return u <= lds_last ? lds[u] : all[__builtin_elementwise_min(u, all_last)];
    }
};

// LDS: entries 0 .. min(lut_max, maxv) of the gamma table (at most 4096:
// bit depths up to 12) staged once per workgroup.
template <int OUT16, bool LDS>
__global__ __launch_bounds__(256) void k_rgb_epilogue(uint8_t* __restrict__ rgb, uint64_t nchunks, uint64_t npx,
                                                      Epilogue e) {
    constexpr int PXC = OUT16 ? 8 : 16;
    __shared__ uint16_t s_lut[LDS ? 4096 : 1];
    const uint32_t lds_last = OUT16 ? (uint32_t)e.maxv : 255u;
    if (LDS) {
        for (int i = (int)threadIdx.x; i <= (int)lds_last; i += (int)blockDim.x) s_lut[i] = e.lut[i];
        __syncthreads();
    }
    const LutSplit split{s_lut, LDS ? lds_last : 0u, e.lut, e.lut_max};
    const LutRead whole{e.lut, e.lut_max};
    auto px = [&](int& r, int& g, int& b) {
        if (LDS) epi_px(e, r, g, b, split);
        else epi_px(e, r, g, b, whole);
    };
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t c = id; c < nchunks; c += stride) {
        uint4* p = (uint4*)(rgb + 48u * c);
        const uint4 v0 = p[0], v1 = p[1], v2 = p[2];
        uint32_t wd[12] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w, v2.x, v2.y, v2.z, v2.w};
        int s[3 * PXC];
#pragma unroll
        for (int i = 0; i < 3 * PXC; ++i)
            s[i] = OUT16 ? (int)((wd[i >> 1] >> (16 * (i & 1))) & 0xFFFFu) : (int)((wd[i >> 2] >> (8 * (i & 3))) & 0xFFu);
#pragma unroll
        for (int j = 0; j < PXC; ++j) px(s[3 * j], s[3 * j + 1], s[3 * j + 2]);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            if (OUT16) wd[k] = (uint32_t)s[2 * k] | ((uint32_t)s[2 * k + 1] << 16);
            else
                wd[k] = (uint32_t)s[4 * k] | ((uint32_t)s[4 * k + 1] << 8) | ((uint32_t)s[4 * k + 2] << 16) |
                        ((uint32_t)s[4 * k + 3] << 24);
        }
        p[0] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
        p[1] = make_uint4(wd[4], wd[5], wd[6], wd[7]);
        p[2] = make_uint4(wd[8], wd[9], wd[10], wd[11]);
    }
    // pixels after the last whole chunk (all of them when the buffer is not 16-byte aligned)
    for (uint64_t q = nchunks * PXC + id; q < npx; q += stride) {
        if (OUT16) {
            uint16_t* t = (uint16_t*)rgb + 3u * q;
            int r = t[0], g = t[1], b = t[2];
            px(r, g, b);
            t[0] = (uint16_t)r; t[1] = (uint16_t)g; t[2] = (uint16_t)b;
        } else {
            uint8_t* t = rgb + 3u * q;
            int r = t[0], g = t[1], b = t[2];
            px(r, g, b);
            t[0] = (uint8_t)r; t[1] = (uint8_t)g; t[2] = (uint8_t)b;
        }
    }
}

}  // namespace

hipError_t launch_rgb_epilogue(uint8_t* rgb, uint64_t npx, int out16, const Epilogue& e, hipStream_t st) {
    if (npx == 0 || e.flags == 0) return hipSuccess;
    const uint64_t pxc = out16 ? 8u : 16u;
    const uint64_t nchunks = ((uintptr_t)rgb & 15u) ? 0u : npx / pxc;
    const uint64_t work = nchunks ? nchunks : npx;
    const uint64_t blocks = (work + 255u) / 256u;
    const dim3 grid((unsigned)(blocks < 16384u ? blocks : 16384u)), block(256);
    const bool lds = (e.flags & EPI_GAMMA) && (out16 ? (uint32_t)e.maxv : 255u) < 4096u;
    if (out16) {
        if (lds) hipLaunchKernelGGL((k_rgb_epilogue<1, true>), grid, block, 0, st, rgb, nchunks, npx, e);
        else hipLaunchKernelGGL((k_rgb_epilogue<1, false>), grid, block, 0, st, rgb, nchunks, npx, e);
    } else {
        if (lds) hipLaunchKernelGGL((k_rgb_epilogue<0, true>), grid, block, 0, st, rgb, nchunks, npx, e);
        else hipLaunchKernelGGL((k_rgb_epilogue<0, false>), grid, block, 0, st, rgb, nchunks, npx, e);
    }
    return hipGetLastError();
}

}  // namespace imop_hip
