// epilogue.hpp -- the integer-domain RGB stages that follow the demosaic,
// applied per output pixel: manual white balance
// (src/white_balance/manual.cpp:8-41), the linear 3x3 colour matrix
// (src/ccm/linear_3x3.cpp:9-46) and gamma (src/tone/gamma.cpp:10-34), in
// that order.
//
// Each stage rounds and clamps to the frame's integer range before the next,
// as the reference's separate passes do.  Float arithmetic keeps the
// reference build's operation order with the multiply-adds GCC fuses there
// (-ffp-contract=fast for C++): fmaf(v, gain, 0.5f) for white balance,
// fmaf(m2, b, fmaf(m0, r, m1 * g)) then + 0.5f for the matrix.  Gamma is a
// table over every storable value (256 for u8 RGB, 65536 for u16), built on
// the host with the same powf the reference calls: the stage is a function
// of the integer value alone, and values above the bit depth's maximum (u8
// frames below 8 bits) are mapped as the reference maps them.
#pragma once

#include "device_common.hpp"

namespace imop_hip {

// detail::clamp_val(static_cast<int>(v), maxv) as the reference's x86-64
// build computes it: cvttss2si truncates, and gives INT_MIN (clamped to 0)
// for NaN and for values of 2^31 or more, where the GPU's conversion would
// saturate instead.
__device__ __forceinline__ int epi_round_clamp(float v, int maxv) {
    if (!(v < 2147483648.0f)) return 0;
    return imin(imax((int)fmaxf(v, -1.0f), 0), maxv);
}

__device__ __forceinline__ void epi_px(const Epilogue& e, int& r, int& g, int& b) {
#pragma clang fp contract(off)
    if (e.flags & EPI_WB) {
        r = epi_round_clamp(fmaf((float)r, e.gain[0], 0.5f), e.maxv);
        g = epi_round_clamp(fmaf((float)g, e.gain[1], 0.5f), e.maxv);
        b = epi_round_clamp(fmaf((float)b, e.gain[2], 0.5f), e.maxv);
    }
    if (e.flags & EPI_CCM) {
        const float fr = (float)r, fg = (float)g, fb = (float)b;
        const float ro = fmaf(e.m[2], fb, fmaf(e.m[0], fr, e.m[1] * fg));
        const float go = fmaf(e.m[5], fb, fmaf(e.m[3], fr, e.m[4] * fg));
        const float bo = fmaf(e.m[8], fb, fmaf(e.m[6], fr, e.m[7] * fg));
        r = epi_round_clamp(ro + 0.5f, e.maxv);
        g = epi_round_clamp(go + 0.5f, e.maxv);
        b = epi_round_clamp(bo + 0.5f, e.maxv);
    }
    if (e.flags & EPI_GAMMA) {  // indices kept inside the table whatever the value
        r = e.lut[__builtin_elementwise_min((uint32_t)r, e.lut_max)];
        g = e.lut[__builtin_elementwise_min((uint32_t)g, e.lut_max)];
        b = e.lut[__builtin_elementwise_min((uint32_t)b, e.lut_max)];
    }
}

// Standalone form over an interleaved RGB image (u8, or u16 when out16).
hipError_t launch_rgb_epilogue(uint8_t* rgb, uint64_t npx, int out16, const Epilogue& e, hipStream_t st);

}  // namespace imop_hip
