// ref_isp_shim.cpp -- C entry points over the reference's own CPU white
// balance, colour matrix, gamma, lens-shading and adaptive defect-correction
// stages and the raw-data analyzer, so tests can call the compiled
// reference through ctypes.  TEST INFRASTRUCTURE ONLY: built by
// oracle/build_ref.sh together with the reference sources under
// /root/reference into oracle/_ref/libimop_ref_isp.so, used to pin
// isp_oracle.c.
#include <cstdint>

#include "ccm/algorithms.hpp"
#include "defect_correct/algorithms.hpp"
#include "imop/analyzer.hpp"
#include "lens_shading/algorithms.hpp"
#include "tone/algorithms.hpp"
#include "white_balance/algorithms.hpp"

#define REF_EXPORT extern "C" __attribute__((visibility("default")))

// white_balance::process_manual_wb, src/white_balance/manual.cpp:8-41
REF_EXPORT int ref_wb_manual(const uint8_t* in, uint8_t* out, int w, int h, int bd, const float* gains) {
    white_balance::WBCoefficients c;
    c.r_gain = gains[0];
    c.g_gain = gains[1];
    c.b_gain = gains[2];
    return static_cast<int>(white_balance::process_manual_wb(in, out, w, h, 3, bd, 0.0f, c));
}

// ccm::process_linear_3x3, src/ccm/linear_3x3.cpp:9-46
REF_EXPORT int ref_ccm_linear(const uint8_t* in, uint8_t* out, int w, int h, int bd, const float* m) {
    ccm::CCMatrix3x3 mat;
    for (int k = 0; k < 9; ++k) mat.m[k] = m[k];
    return static_cast<int>(ccm::process_linear_3x3(in, out, w, h, 3, bd, &mat));
}

// tone::process_gamma, src/tone/gamma.cpp:10-34
REF_EXPORT int ref_gamma(const uint8_t* in, uint8_t* out, int w, int h, int bd, float gamma) {
    tone::ToneParams p;
    p.gamma = gamma;
    return static_cast<int>(tone::process_gamma(in, out, w, h, 3, bd, p));
}

// lens_shading::process_polynomial, src/lens_shading/polynomial.cpp:11-52
// (in place; coef = {a2, a4, a6} for R, Gr, Gb, B)
REF_EXPORT int ref_lens_shading_poly(uint8_t* data, int w, int h, int pattern, int bd, const float* coef, float cx,
                                     float cy) {
    lens_shading::LensShadingParams p;
    lens_shading::ShadingPolynomial* c[4] = {&p.r_coef, &p.gr_coef, &p.gb_coef, &p.b_coef};
    for (int k = 0; k < 4; ++k) {
        c[k]->a2 = coef[3 * k];
        c[k]->a4 = coef[3 * k + 1];
        c[k]->a6 = coef[3 * k + 2];
    }
    p.center_x = cx;
    p.center_y = cy;
    return static_cast<int>(lens_shading::process_polynomial(data, w, h, static_cast<imop::BayerPattern>(pattern), bd, p));
}

// defect_correct::process_adaptive, src/defect_correct/adaptive.cpp:8-61 (in place)
REF_EXPORT int ref_defect_adaptive(uint8_t* data, int w, int h, int pattern, int bd, float threshold) {
    defect_correct::DefectCorrectParams p;
    p.threshold = threshold;
    return static_cast<int>(
        defect_correct::process_adaptive(data, w, h, static_cast<imop::BayerPattern>(pattern), bd, p));
}

// imop::analyze_data, src/analyzer.cpp:151-214: out = {detected_bit_depth,
// suggested_width, suggested_height, pixel_count, max_value, min_value,
// is_likely_16bit, is_packed, n_dims, then up to 8 (w, h) pairs}.
REF_EXPORT void ref_analyze_data(const uint8_t* data, uint64_t bytes, int* out) {
    const imop::DataInfo d = imop::analyze_data(data, static_cast<size_t>(bytes));
    out[0] = d.detected_bit_depth;
    out[1] = d.suggested_width;
    out[2] = d.suggested_height;
    out[3] = d.pixel_count;
    out[4] = d.max_value;
    out[5] = d.min_value;
    out[6] = d.is_likely_16bit ? 1 : 0;
    out[7] = d.is_packed ? 1 : 0;
    const int n = static_cast<int>(d.possible_dimensions.size() < 8 ? d.possible_dimensions.size() : 8);
    out[8] = n;
    for (int k = 0; k < n; ++k) {
        out[9 + 2 * k] = d.possible_dimensions[k].first;
        out[10 + 2 * k] = d.possible_dimensions[k].second;
    }
}

// imop::guess_pattern, src/analyzer.cpp:216-321
REF_EXPORT int ref_guess_pattern(const uint8_t* data, int w, int h, int bd, int packed) {
    return static_cast<int>(imop::guess_pattern(data, w, h, bd, packed != 0));
}
