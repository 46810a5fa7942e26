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
// table over every storable value (256 entries for u8 RGB, 65536 for u16),
// built on the host with the same powf the reference calls: the stage is a
// function of the integer value alone, values above the bit depth's maximum
// included (u8 frames below 8 bits, or any u16 image handed to the
// standalone pass), which map to the maximum or, where c * max reaches 2^31,
// to 0 as in the reference.
#pragma once

#include "device_common.hpp"

namespace imop_hip {

// detail::clamp_val(static_cast<int>(v), maxv) as the reference's x86-64
// build computes it: cvttss2si truncates, and gives INT_MIN (clamped to 0)
// for NaN and for values of 2^31 or more, where the GPU's conversion would
// saturate instead.  BOUNDED (the host has checked that no stage value can
// reach 2^31 and no parameter is NaN or infinite): clamping in float first,
// then truncating, gives the same integer in two instructions.
template <bool BOUNDED>
__device__ __forceinline__ int epi_round_clamp(float v, int maxv) {
    if (BOUNDED) return (int)__builtin_amdgcn_fmed3f(v, 0.0f, (float)maxv);
    if (!(v < 2147483648.0f)) return 0;
    return imin(imax((int)fmaxf(v, -1.0f), 0), maxv);
}

// One pixel through the stages; `look(v)` is the gamma table lookup.
// BOUNDED keeps the values in float between the stages: a stage's result
// (int)med3(v, 0, max) enters the next as its float, which is
// truncf(med3(v, 0, max)) exactly (integers up to 65535).
template <bool BOUNDED, typename Look>
__device__ __forceinline__ void epi_px_t(const Epilogue& e, int& r, int& g, int& b, Look look) {
#pragma clang fp contract(off)
    if (BOUNDED && (e.flags & (EPI_WB | EPI_CCM))) {
        const float mx = (float)e.maxv;
        float fr = (float)r, fg = (float)g, fb = (float)b;
        if (e.flags & EPI_WB) {
            fr = truncf(__builtin_amdgcn_fmed3f(fmaf(fr, e.gain[0], 0.5f), 0.0f, mx));
            fg = truncf(__builtin_amdgcn_fmed3f(fmaf(fg, e.gain[1], 0.5f), 0.0f, mx));
            fb = truncf(__builtin_amdgcn_fmed3f(fmaf(fb, e.gain[2], 0.5f), 0.0f, mx));
        }
        if (e.flags & EPI_CCM) {
            const float ro = fmaf(e.m[2], fb, fmaf(e.m[0], fr, e.m[1] * fg));
            const float go = fmaf(e.m[5], fb, fmaf(e.m[3], fr, e.m[4] * fg));
            const float bo = fmaf(e.m[8], fb, fmaf(e.m[6], fr, e.m[7] * fg));
            fr = __builtin_amdgcn_fmed3f(ro + 0.5f, 0.0f, mx);
            fg = __builtin_amdgcn_fmed3f(go + 0.5f, 0.0f, mx);
            fb = __builtin_amdgcn_fmed3f(bo + 0.5f, 0.0f, mx);
        }
        r = (int)fr;
        g = (int)fg;
        b = (int)fb;
        if (e.flags & EPI_GAMMA) {
            r = look(r);
            g = look(g);
            b = look(b);
        }
        return;
    }
    if (e.flags & EPI_WB) {
        r = epi_round_clamp<BOUNDED>(fmaf((float)r, e.gain[0], 0.5f), e.maxv);
        g = epi_round_clamp<BOUNDED>(fmaf((float)g, e.gain[1], 0.5f), e.maxv);
        b = epi_round_clamp<BOUNDED>(fmaf((float)b, e.gain[2], 0.5f), e.maxv);
    }
    if (e.flags & EPI_CCM) {
        const float fr = (float)r, fg = (float)g, fb = (float)b;
        const float ro = fmaf(e.m[2], fb, fmaf(e.m[0], fr, e.m[1] * fg));
        const float go = fmaf(e.m[5], fb, fmaf(e.m[3], fr, e.m[4] * fg));
        const float bo = fmaf(e.m[8], fb, fmaf(e.m[6], fr, e.m[7] * fg));
        r = epi_round_clamp<BOUNDED>(ro + 0.5f, e.maxv);
        g = epi_round_clamp<BOUNDED>(go + 0.5f, e.maxv);
        b = epi_round_clamp<BOUNDED>(bo + 0.5f, e.maxv);
    }
    if (e.flags & EPI_GAMMA) {
        r = look(r);
        g = look(g);
        b = look(b);
    }
}

template <typename Look>
__device__ __forceinline__ void epi_px(const Epilogue& e, int& r, int& g, int& b, Look look) {
    if (e.bounded) epi_px_t<true>(e, r, g, b, look);
    else epi_px_t<false>(e, r, g, b, look);
}

// Table lookups: entries 0 .. n - 1 of `lut` (LDS or global memory), the
// index clamped into the table whatever the value.
struct LutRead {
    const uint16_t* lut;
    uint32_t last;
    __device__ __forceinline__ int operator()(int v) const {
        return lut[__builtin_elementwise_min((uint32_t)v, last)];
    }
};

// Standalone form over an interleaved RGB image (u8, or u16 when out16).
hipError_t launch_rgb_epilogue(uint8_t* rgb, uint64_t npx, int out16, const Epilogue& e, hipStream_t st);

// Polynomial lens shading in place on an unpacked Bayer frame (u8 for
// bd <= 8, else u16); coef = {a2, a4, a6} for R, Gr, Gb, B.
hipError_t launch_lens_shading(uint8_t* data, int w, int h, int pattern, int bd, const float coef[12], float center_x,
                               float center_y, hipStream_t st);

// The lens-shading constants of one frame (the reference's per-call cx, cy,
// max_r; coefficient set per CFA parity); on = 1.
LensParams make_lens_params(int w, int h, int pattern, const float coef[12], float center_x, float center_y);

// Bayer prologue into an unpacked frame (u8 for bd <= 8, else u16): every
// sample decoded from the input layout of f, black level when f.bl_on, then
// the lens-shading gain -- unpack -> process_black_level ->
// process_polynomial as the reference runs them, for the demosaic paths that
// cannot take the gain in their decode.
hipError_t launch_bayer_prologue(const FrameParams& f, const LensParams& lens, uint8_t* out, hipStream_t st);

// Adaptive defect correction (defect_kernels.hip): the copy of the input
// into `dst` (also clearing ctl[0..3]) and one fixed-point pass; tiles of
// 128 x 32 pixels, `flags` holds 2 bytes per tile.
hipError_t launch_defect_copy(const uint8_t* src, uint8_t* dst, uint64_t nbytes, int* ctl, hipStream_t st);
hipError_t launch_defect_pass(uint8_t* data, const uint8_t* orig, uint8_t* flags, int* ctl, int w, int h, int bd,
                              int thresh, int pass, hipStream_t st);
hipError_t launch_defect_status(const int* ctl, int passes, int* status, hipStream_t st);
int defect_tiles_x(int w);
int defect_tiles_y(int h);

}  // namespace imop_hip
