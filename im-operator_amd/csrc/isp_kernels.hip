// isp_kernels.hip -- in-place passes around the demosaic:
//  * the RGB epilogue (white balance, colour matrix, gamma; epilogue.hpp) over
//    an interleaved RGB image.  The demosaic paths that cannot apply it in
//    registers (PRISM, the generic LDS-tiled kernels) run it after the
//    demosaic; it is also the standalone imop_hip_rgb_epilogue_apply;
//  * the polynomial lens-shading correction over an unpacked Bayer frame.
#include <cmath>

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

// Polynomial lens shading (lens_shading::process_polynomial,
// src/lens_shading/polynomial.cpp:11-52) in place, four samples of a row per
// thread.  cx, cy and max_r are the reference's per-call values, computed on
// the host with the same float operations.  Per sample, in the reference
// build's order with its fused multiply-adds: r = sqrt(fma(dx, dx, dy * dy))
// / max_r (correctly rounded), gain = fma(a6, r^6, fma(a4, r^4, fma(a2, r^2,
// 1))), v' = clamp_val(int(fma(v, gain, 0.5)), max) with the x86 conversion.
struct LensShading {
    float coef[4][3];   // {a2, a4, a6} for R, Gr, Gb, B
    float cx, cy, max_r;
    int32_t maxv;
    int32_t r_row, r_col, b_row, b_col;  // PatternOffsets
};

template <int U16>
__global__ __launch_bounds__(256) void k_lens_shading(uint8_t* __restrict__ data, int w, int h, LensShading p) {
#pragma clang fp contract(off)
    const int x0 = 4 * (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (x0 >= w) return;
    for (int y = (int)blockIdx.y; y < h; y += (int)gridDim.y) {
        const float dy = (float)y - p.cy;
        const float dyy = dy * dy;
        const bool r_row = (y & 1) == p.r_row, b_row = (y & 1) == p.b_row;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int x = x0 + k;
            if (x >= w) break;
            const float dx = (float)x - p.cx;
            const float r = sqrtf(fmaf(dx, dx, dyy)) / p.max_r;  // both correctly rounded (hipcc default)
            // lens_shading::detail::bayer_color: 0 R, 1 Gr, 2 Gb, 3 B
            const int c = (r_row && (x & 1) == p.r_col) ? 0 : (b_row && (x & 1) == p.b_col) ? 3 : r_row ? 1 : 2;
            const float r2 = r * r, r4 = r2 * r2, r6 = r4 * r2;
            const float gain = fmaf(p.coef[c][2], r6, fmaf(p.coef[c][1], r4, fmaf(p.coef[c][0], r2, 1.0f)));
            const uint64_t i = (uint64_t)y * (uint64_t)w + (uint64_t)x;
            const int v = U16 ? ((const uint16_t*)data)[i] : data[i];
            const int o = epi_round_clamp<false>(fmaf((float)v, gain, 0.5f), p.maxv);
            if (U16) ((uint16_t*)data)[i] = (uint16_t)o;
            else data[i] = (uint8_t)o;
        }
    }
}

// Bayer prologue: decode (no MSB alignment: the lens stage reads the
// stored value), black level, lens gain; one sample per thread.
__global__ __launch_bounds__(256) void k_bayer_prologue(FrameParams f, LensParams lens, uint8_t* __restrict__ out) {
#pragma clang fp contract(off)
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)blockIdx.y;
    if (x >= f.w) return;
    for (int yy = y; yy < f.h; yy += (int)gridDim.y) {
        const uint64_t i = (uint64_t)yy * (uint64_t)f.w + (uint64_t)x;
        int v;
        switch (f.fmt) {
        case FMT_U8: v = f.in[i]; break;
        case FMT_U16: v = (int)((const uint16_t*)f.in)[i]; break;
        case FMT_P12: v = decode_p12(f.in, i, f.in_bytes); break;
        case FMT_P10: v = decode_p10(f.in, i, f.in_bytes); break;
        case FMT_G10: v = decode_g10(f.in, i, f.in_bytes); break;
        default: v = decode_pgen(f.in, i, f.in_bytes, f.bd); break;
        }
        const int par = ((yy & 1) << 1) | (x & 1);
        if (f.bl_on) v = iclamp((int)((float)v - f.bl_q[par]), 0, f.maxv);
        const float dy = (float)yy - lens.cy;
        const float dyy = dy * dy;
        const float dx = (float)x - lens.cx;
        const float r = sqrtf(fmaf(dx, dx, dyy)) / lens.max_r;
        const float r2 = r * r, r4 = r2 * r2, r6 = r4 * r2;
        const float gain = fmaf(lens.pc[par][2], r6, fmaf(lens.pc[par][1], r4, fmaf(lens.pc[par][0], r2, 1.0f)));
        const int o = epi_round_clamp<false>(fmaf((float)v, gain, 0.5f), f.maxv);
        if (f.bd > 8) ((uint16_t*)out)[i] = (uint16_t)o;
        else out[i] = (uint8_t)o;
    }
}

}  // namespace

LensParams make_lens_params(int w, int h, int pattern, const float coef[12], float center_x, float center_y) {
#pragma clang fp contract(off)
    LensParams p{};
    // the reference's per-call constants (polynomial.cpp:26-28), GCC's contraction kept
    p.cx = center_x * (float)w;
    p.cy = center_y * (float)h;
    p.max_r = sqrtf(fmaf(p.cx, p.cx, p.cy * p.cy));
    p.on = 1;
    // lens_shading::detail::bayer_color (common.hpp:28-33) per parity
    static const int off[4][4] = {{0, 0, 1, 1}, {1, 1, 0, 0}, {0, 1, 1, 0}, {1, 0, 0, 1}};  // r_row r_col b_row b_col
    const int* o = off[(pattern >= 0 && pattern < 4) ? pattern : 0];
    for (int py = 0; py < 2; ++py)
        for (int px = 0; px < 2; ++px) {
            const int c = (py == o[0] && px == o[1]) ? 0 : (py == o[2] && px == o[3]) ? 3 : py == o[0] ? 1 : 2;
            for (int k = 0; k < 3; ++k) p.pc[py * 2 + px][k] = coef[3 * c + k];
        }
    // 1 / max_r rounded to nearest: the double quotient's float, or the
    // neighbour whose product with max_r is closer to 1 (products exact in double)
    if (p.max_r > 0.0f && std::isfinite(p.max_r)) {
        const float y = (float)(1.0 / (double)p.max_r);
        float best = y;
        double err = std::fabs((double)p.max_r * (double)y - 1.0);
        for (const float c : {std::nextafter(y, 0.0f), std::nextafter(y, INFINITY)}) {
            const double e = std::fabs((double)p.max_r * (double)c - 1.0);
            if (e < err) { err = e; best = c; }
        }
        p.inv_max_r = best;
    }
    // Range of the register-resident kernel's sequences.  A centre
    // coordinate of 0 or of magnitude in [2^-20, 2^40] keeps every nonzero
    // sum of squares at or above 2^-86 (no scaling for the sqrt) and every
    // quotient r normal; the gain bound keeps v * gain + 0.5 below 2^31.
    auto centre_ok = [](float c) {
        const float a = std::fabs(c);
        return a == 0.0f || (a >= 0x1p-20f && a <= 0x1p40f);
    };
    bool ok = centre_ok(p.cx) && centre_ok(p.cy) && p.max_r > 0.0f && std::isfinite(p.max_r);
    if (ok) {
        const double dx = std::fmax(std::fabs((double)p.cx), std::fabs((double)(w - 1) - (double)p.cx));
        const double dy = std::fmax(std::fabs((double)p.cy), std::fabs((double)(h - 1) - (double)p.cy));
        const double R = std::sqrt(dx * dx + dy * dy) / (double)p.max_r * (1.0 + 1e-5);
        const double R2 = R * R;
        double g = 0.0;
        for (int c = 0; c < 4; ++c) {
            const double gc = 1.0 + std::fabs((double)coef[3 * c]) * R2 + std::fabs((double)coef[3 * c + 1]) * R2 * R2 +
                              std::fabs((double)coef[3 * c + 2]) * R2 * R2 * R2;
            g = std::fmax(g, gc);
        }
        ok = std::isfinite(g) && 65535.0 * g * (1.0 + 1e-4) + 2.0 < 2147483648.0;
    }
    p.fused_ok = ok ? 1 : 0;
    return p;
}

hipError_t launch_bayer_prologue(const FrameParams& f, const LensParams& lens, uint8_t* out, hipStream_t st) {
    const dim3 grid((unsigned)((f.w + 255) / 256), (unsigned)(f.h < 65535 ? f.h : 65535)), block(256);
    hipLaunchKernelGGL(k_bayer_prologue, grid, block, 0, st, f, lens, out);
    return hipGetLastError();
}

hipError_t launch_lens_shading(uint8_t* data, int w, int h, int pattern, int bd, const float coef[12], float center_x,
                               float center_y, hipStream_t st) {
#pragma clang fp contract(off)
    LensShading p;
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 3; ++k) p.coef[c][k] = coef[3 * c + k];
    // the reference's per-call constants (polynomial.cpp:26-28), same operations
    p.cx = center_x * (float)w;
    p.cy = center_y * (float)h;
    p.max_r = sqrtf(fmaf(p.cx, p.cx, p.cy * p.cy));
    p.maxv = bd <= 0 ? 255 : bd > 16 ? 65535 : (1 << bd) - 1;  // lens_shading::detail::safe_max_val
    static const int off[4][4] = {{0, 0, 1, 1}, {1, 1, 0, 0}, {0, 1, 1, 0}, {1, 0, 0, 1}};  // PatternOffsets
    const int* o = off[(pattern >= 0 && pattern < 4) ? pattern : 0];
    p.r_row = o[0]; p.r_col = o[1]; p.b_row = o[2]; p.b_col = o[3];
    const unsigned bx = (unsigned)((w + 1023) / 1024);
    const dim3 grid(bx, (unsigned)(h < 65535 ? h : 65535)), block(256);
    if (bd > 8) hipLaunchKernelGGL((k_lens_shading<1>), grid, block, 0, st, data, w, h, p);
    else hipLaunchKernelGGL((k_lens_shading<0>), grid, block, 0, st, data, w, h, p);
    return hipGetLastError();
}

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
