/*
 * isp_oracle.c -- CPU restatement of im-operator's integer-domain RGB stages
 * that follow the demosaic (manual white balance, the linear 3x3 colour
 * matrix, gamma) and of the polynomial lens-shading correction that runs on
 * the Bayer frame before it.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP path is compared
 * against; it is never linked into, loaded by, or called from the product
 * library (im-operator_amd/).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * What it restates:
 *   - manual white balance     src/white_balance/manual.cpp:8-41
 *   - linear 3x3 colour matrix src/ccm/linear_3x3.cpp:9-46
 *   - gamma                    src/tone/gamma.cpp:10-34
 *   - polynomial lens shading  src/lens_shading/polynomial.cpp:11-52,
 *                              src/lens_shading/common.hpp:12-35
 *   - adaptive defect correct. src/defect_correct/adaptive.cpp:8-61,
 *                              src/defect_correct/common.hpp:12-16
 *   - helpers                  src/white_balance/common.hpp:13-47 (clamp_val,
 *                              safe_max_val, read_pixel, write_pixel; the ccm
 *                              and tone copies are the same)
 *
 * Floating point keeps the reference's operation order.  This file is
 * compiled with -ffp-contract=off; the multiply-adds GCC fuses in the
 * reference build (-O3 -mavx2 -mfma, C++: -ffp-contract=fast by default) are
 * explicit fmaf() calls: `r * gain + 0.5f` -> fmaf(r, gain, 0.5f);
 * `m0*r + m1*g + m2*b` -> fmaf(m2, b, fmaf(m0, r, m1 * g)) (GCC fuses the
 * first product of each sum into the addition); `corrected * mv + 0.5f` ->
 * fmaf; lens shading: `cx*cx + cy*cy` -> fmaf(cx, cx, cy * cy), the gain
 * polynomial -> fmaf(a6, r6, fmaf(a4, r4, fmaf(a2, r2, 1.0f))),
 * `v * gain + 0.5f` -> fmaf.  static_cast<int> of a float is cvttss2si on x86-64: truncation,
 * and INT_MIN for NaN or anything outside [-2^31, 2^31).
 *
 * Pinning: tests/test_isp_oracle.py checks every function against the
 * compiled reference stages (oracle/_ref/libimop_ref_isp.so, built from the
 * reference sources by oracle/build_ref.sh) on random images with ordinary
 * and extreme parameters, against closed forms (unit gains, identity matrix
 * and gamma 1 leave an image unchanged; a permutation matrix permutes the
 * channels) and against the reference's lens-shading and defect-correction
 * unit tests (tests/unit/test_lens_shading.cpp:24-84,
 * tests/unit/test_defect_correct.cpp:8-72).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

/* static_cast<int>(float) as compiled for x86-64 (cvttss2si) */
static int cvt_x86(float v) {
    return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : INT32_MIN;
}

/* detail::clamp_val: std::max(0, std::min(v, max_val)) */
static int clamp_val(int v, int m) {
    const int lo = v < m ? v : m;
    return lo > 0 ? lo : 0;
}

/* detail::safe_max_val */
int isp_max_val(int bd) {
    if (bd <= 0) return 255;
    if (bd >= 16) return 65535;
    return (1 << bd) - 1;
}

static int rd(const uint8_t *d, size_t i, int bd) { return bd <= 8 ? d[i] : ((const uint16_t *)d)[i]; }
static void wr(uint8_t *d, size_t i, int bd, int v) {
    if (bd <= 8) d[i] = (uint8_t)v;
    else ((uint16_t *)d)[i] = (uint16_t)v;
}

/* white_balance::process_manual_wb (3 channels) */
void isp_wb_manual(const uint8_t *in, uint8_t *out, int w, int h, int bd, const float gains[3]) {
    const int mv = isp_max_val(bd);
    const size_t n = (size_t)w * (size_t)h;
    for (size_t p = 0; p < n; ++p)
        for (int c = 0; c < 3; ++c)
            wr(out, 3 * p + c, bd, clamp_val(cvt_x86(fmaf((float)rd(in, 3 * p + c, bd), gains[c], 0.5f)), mv));
}

/* ccm::process_linear_3x3 (3 channels; m row-major, out = M * in) */
void isp_ccm_linear(const uint8_t *in, uint8_t *out, int w, int h, int bd, const float m[9]) {
    const int mv = isp_max_val(bd);
    const size_t n = (size_t)w * (size_t)h;
    for (size_t p = 0; p < n; ++p) {
        const float r = (float)rd(in, 3 * p, bd), g = (float)rd(in, 3 * p + 1, bd), b = (float)rd(in, 3 * p + 2, bd);
        float o[3];
        for (int c = 0; c < 3; ++c) o[c] = fmaf(m[3 * c + 2], b, fmaf(m[3 * c], r, m[3 * c + 1] * g));
        for (int c = 0; c < 3; ++c) wr(out, 3 * p + c, bd, clamp_val(cvt_x86(o[c] + 0.5f), mv));
    }
}

/* tone::process_gamma: the exponent is clamped to [0.1, 10] with
 * std::max(0.1f, std::min(10.0f, gamma)) (NaN -> 10) */
float isp_gamma_clamp(float gamma) {
    const float lo = (gamma < 10.0f) ? gamma : 10.0f;
    return (0.1f < lo) ? lo : 0.1f;
}

void isp_gamma(const uint8_t *in, uint8_t *out, int w, int h, int bd, float gamma) {
    const int mvi = isp_max_val(bd);
    const float mv = (float)mvi;
    const float inv_gamma = 1.0f / isp_gamma_clamp(gamma);
    const size_t n = (size_t)w * (size_t)h * 3u;
    for (size_t i = 0; i < n; ++i) {
        const float norm = (float)rd(in, i, bd) / mv;
        const float corrected = powf(norm, inv_gamma);
        wr(out, i, bd, clamp_val(cvt_x86(fmaf(corrected, mv, 0.5f)), mvi));
    }
}

/* The three stages in pipeline order, in place on an interleaved RGB image
 * (flags: 1 white balance, 2 colour matrix, 4 gamma). */
void isp_epilogue(uint8_t *rgb, int w, int h, int bd, int flags, const float gains[3], const float m[9],
                  float gamma) {
    if (flags & 1) isp_wb_manual(rgb, rgb, w, h, bd, gains);
    if (flags & 2) isp_ccm_linear(rgb, rgb, w, h, bd, m);
    if (flags & 4) isp_gamma(rgb, rgb, w, h, bd, gamma);
}

/* lens_shading::detail::safe_max_val (src/lens_shading/common.hpp:12-16) */
static int ls_max_val(int bd) {
    if (bd <= 0) return 255;
    if (bd > 16) return 65535;
    return (1 << bd) - 1;
}

/* lens_shading::detail::bayer_color: 0 R, 1 Gr (G in the R row), 2 Gb, 3 B
 * (src/lens_shading/common.hpp:28-33, PatternOffsets of include/imop/types.hpp:88-101) */
static int ls_color(int y, int x, int pattern) {
    static const int off[4][4] = {{0, 0, 1, 1}, {1, 1, 0, 0}, {0, 1, 1, 0}, {1, 0, 0, 1}}; /* r_row r_col b_row b_col */
    const int *o = off[(pattern >= 0 && pattern < 4) ? pattern : 0];
    if ((y & 1) == o[0] && (x & 1) == o[1]) return 0;
    if ((y & 1) == o[2] && (x & 1) == o[3]) return 3;
    if ((y & 1) == o[0]) return 1;
    return 2;
}

/* lens_shading::process_polynomial, in place on an unpacked Bayer frame
 * (u8 for bd <= 8, else u16).  coef: 4 x {a2, a4, a6} for R, Gr, Gb, B;
 * center_x / center_y normalised to the frame. */
void isp_lens_shading_poly(uint8_t *data, int w, int h, int pattern, int bd, const float coef[12], float center_x,
                           float center_y) {
    const int mv = ls_max_val(bd);
    const float cx = center_x * (float)w, cy = center_y * (float)h;
    const float max_r = sqrtf(fmaf(cx, cx, cy * cy));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float dx = (float)x - cx, dy = (float)y - cy;
            const float r = sqrtf(fmaf(dx, dx, dy * dy)) / max_r;
            const float *c = coef + 3 * ls_color(y, x, pattern);
            const float r2 = r * r, r4 = r2 * r2, r6 = r4 * r2;
            const float gain = fmaf(c[2], r6, fmaf(c[1], r4, fmaf(c[0], r2, 1.0f)));
            const size_t i = (size_t)y * (size_t)w + (size_t)x;
            wr(data, i, bd, clamp_val(cvt_x86(fmaf((float)rd(data, i, bd), gain, 0.5f)), mv));
        }
}

/* defect_correct::process_adaptive, in place on an unpacked Bayer frame (u8
 * for bd <= 8, else u16): in raster order over [2, w-2) x [2, h-2), each
 * pixel is compared with the median (element n/2 of the sorted values) of
 * its same-colour neighbours at distance 2 that lie in that region, as they
 * are at that moment (the ones above and to the left already corrected), and
 * replaced by it when they differ by more than threshold * max.  Returns 5
 * (ImageTooSmall) below 5 x 5. */
int isp_defect_adaptive(uint8_t *data, int w, int h, int bd, float threshold) {
    if (w < 5 || h < 5) return 5;
    const int mv = bd <= 0 ? 255 : bd > 16 ? 65535 : (int)((1u << bd) - 1u);
    /* std::max(0.05f, std::min(1.0f, threshold)) */
    const float t_lo = (threshold < 1.0f) ? threshold : 1.0f;
    const float t = (0.05f < t_lo) ? t_lo : 0.05f;
    const int abs_thresh = (int)(t * (float)mv);
    for (int y = 2; y < h - 2; ++y)
        for (int x = 2; x < w - 2; ++x) {
            const int center = rd(data, (size_t)y * w + x, bd);
            int nb[8], n = 0;
            for (int dy = -2; dy <= 2; dy += 2)
                for (int dx = -2; dx <= 2; dx += 2) {
                    if (dx == 0 && dy == 0) continue;
                    const int nx = x + dx, ny = y + dy;
                    if (nx >= 2 && nx < w - 2 && ny >= 2 && ny < h - 2) nb[n++] = rd(data, (size_t)ny * w + nx, bd);
                }
            if (n == 0) continue;
            for (int i = 1; i < n; ++i) /* insertion sort: the sorted order std::sort gives */
                for (int j = i; j > 0 && nb[j - 1] > nb[j]; --j) {
                    const int tmp = nb[j]; nb[j] = nb[j - 1]; nb[j - 1] = tmp;
                }
            const int median = nb[n / 2];
            const int diff = center > median ? center - median : median - center;
            if (diff > abs_thresh) wr(data, (size_t)y * w + x, bd, median);
        }
    return 0;
}

/* ------------------------------------------------------------ raw analyzer */
/* match_common_resolution (src/analyzer.cpp:13-25) */
static size_t an_common(size_t px) {
    static const size_t sizes[] = {640 * 480,   720 * 480,   720 * 576,   800 * 600,   1024 * 768,  1280 * 720,
                                   1280 * 800,  1280 * 1024, 1440 * 900,  1920 * 1080, 1920 * 1200, 2048 * 1536,
                                   2560 * 1440, 2560 * 1600, 3840 * 2160, 4096 * 2160, 4096 * 3072, 5120 * 2880,
                                   5320 * 4600, 7680 * 4320, 1006 * 758,  4024 * 3032};
    for (size_t k = 0; k < sizeof sizes / sizeof sizes[0]; ++k)
        if (px == sizes[k]) return px;
    return 0;
}

static unsigned an_u16(const uint8_t *d, size_t i) { return (unsigned)d[2 * i] | ((unsigned)d[2 * i + 1] << 8); }

/* detect_bit_depth (src/analyzer.cpp:27-117) */
int isp_detect_bit_depth(const uint8_t *d, size_t n) {
    if (!d || n == 0) return 0;
    if (an_common(n)) return 8;
    if (n % 2 == 0 && an_common(n / 2)) { /* guess_bit_from_16bit_data */
        const size_t c = n / 2 < 100000 ? n / 2 : 100000;
        unsigned mx = 0;
        for (size_t i = 0; i < c; ++i) if (an_u16(d, i) > mx) mx = an_u16(d, i);
        return mx <= 255 ? 8 : mx <= 1023 ? 10 : mx <= 4095 ? 12 : mx <= 16383 ? 14 : 16;
    }
    if (n >= 8) { /* looks_like_packed */
        const size_t c = n / 2 < 10000 ? n / 2 : 10000;
        int hi = 0, zero = 0;
        for (size_t i = 0; i < c; ++i) {
            const unsigned v = an_u16(d, i);
            if (v >> 8) hi++;
            if (v == 0) zero++;
        }
        if ((double)hi / (double)c > 0.85 && (double)zero / (double)c < 0.01) {
            static const int ds[4] = {10, 12, 14, 16};
            for (int k = 0; k < 4; ++k) {
                const size_t pc = n * 8 / (size_t)ds[k];
                if (an_common(pc) || an_common(pc + 1)) return ds[k];
            }
        }
    }
    if (n % 2 == 0 && n >= 4) {
        const size_t c = n / 2 < 10000 ? n / 2 : 10000;
        int hi = 0;
        unsigned mx = 0;
        for (size_t i = 0; i < c; ++i) {
            const unsigned v = an_u16(d, i);
            if (v >> 8) hi++;
            if (v > mx) mx = v;
        }
        if ((double)hi / (double)c > 0.3 && mx > 255) return 16;
    }
    return 8;
}

/* suggest_dimensions (src/analyzer.cpp:119-149): up to 8 pairs into dims */
int isp_suggest_dimensions(size_t px, int dims[][2], int cap) {
    static const int cw[] = {640, 720, 800, 1024, 1280, 1440, 1920, 2048, 2560, 3840, 4096, 5120, 7680};
    static const int ch[] = {480, 576, 600, 720, 768, 800, 1024, 1080, 1200, 1440, 1536, 2048, 2160, 3072, 4320};
    int n = 0;
    if (px == 0) return 0;
    for (size_t a = 0; a < sizeof cw / sizeof cw[0]; ++a)
        if (px % (size_t)cw[a] == 0) {
            const int h = (int)(px / (size_t)cw[a]);
            for (size_t b = 0; b < sizeof ch / sizeof ch[0]; ++b)
                if (h == ch[b]) { if (n < cap) { dims[n][0] = cw[a]; dims[n][1] = h; } n++; break; }
        }
    if (n == 0) {
        const int s = (int)sqrt((double)px);
        for (int w = s - 100; w <= s + 100; ++w)
            if (w > 0 && px % (size_t)w == 0) {
                const int h = (int)(px / (size_t)w);
                if (w >= 64 && h >= 64) {
                    if (n < cap) { dims[n][0] = w; dims[n][1] = h; }
                    if (++n >= 5) break;
                }
            }
    }
    return n;
}

/* analyze_data (src/analyzer.cpp:151-214): out as ref_analyze_data */
void isp_analyze_data(const uint8_t *d, size_t n, int *out) {
    for (int k = 0; k < 25; ++k) out[k] = 0;
    if (!d || n == 0) return;
    const int bd = isp_detect_bit_depth(d, n);
    int pc, mn, mx = 0, packed = 0;
    if (bd > 8) {
        const size_t p16 = n / 2;
        if (an_common(p16)) {
            pc = (int)p16;
            mn = 65535;
            for (size_t i = 0; i < p16; ++i) {
                const int v = (int)an_u16(d, i);
                if (v < mn) mn = v;
                if (v > mx) mx = v;
            }
        } else {
            packed = 1;
            pc = (int)(n * 8 / (size_t)bd);
            if (!an_common((size_t)pc) && an_common((size_t)pc + 1)) pc++;
            mn = bd >= 16 ? 65535 : (1 << bd) - 1;
            const int sample = pc < 10000 ? pc : 10000;
            for (int i = 0; i < sample; ++i) {
                const int bo = (i * bd) / 8, bs = (i * bd) % 8;
                if ((size_t)bo + 1 < n) {
                    int v = (d[bo] | (d[bo + 1] << 8)) >> bs;
                    v &= (int)((1u << bd) - 1u);
                    if (v < mn) mn = v;
                    if (v > mx) mx = v;
                }
            }
        }
    } else {
        pc = (int)n;
        mn = 255;
        for (size_t i = 0; i < n; ++i) {
            if (d[i] < mn) mn = d[i];
            if (d[i] > mx) mx = d[i];
        }
    }
    out[0] = bd; out[3] = pc; out[4] = mx; out[5] = mn; out[6] = bd > 8; out[7] = packed;
    int dims[8][2];
    const int nd = isp_suggest_dimensions((size_t)pc, dims, 8);
    out[8] = nd < 8 ? nd : 8;
    for (int k = 0; k < out[8]; ++k) { out[9 + 2 * k] = dims[k][0]; out[10 + 2 * k] = dims[k][1]; }
    if (nd > 0) { out[1] = dims[0][0]; out[2] = dims[0][1]; }
}

/* guess_pattern (src/analyzer.cpp:216-321); samples as the demosaic reads
 * them (8-bit bytes, u16 aligned, packed decoded: oracle's orc_packed_pixel
 * semantics are reproduced here for the three layouts) */
static int an_val(const uint8_t *d, int x, int y, int w, int h, int bd, int packed, size_t nb) {
    if (x < 0 || x >= w || y < 0 || y >= h) return 0;
    const size_t i = (size_t)y * (size_t)w + (size_t)x;
    if (bd <= 8) return d[i];
    if (packed) {
        extern int orc_packed_pixel(const uint8_t *, size_t, int, size_t);
        return orc_packed_pixel(d, i, bd, nb);
    }
    unsigned v = an_u16(d, i);
    if (bd < 16 && v > (1u << bd) - 1u) v >>= (16 - bd); /* align_raw_value */
    return (int)v;
}

int isp_guess_pattern(const uint8_t *d, int w, int h, int bd, int packed) {
    if (!d || w < 4 || h < 4) return 0;
    const size_t nb = packed ? ((size_t)w * (size_t)h * (size_t)bd + 7u) / 8u : (size_t)w * h * (bd <= 8 ? 1u : 2u);
    const long long npx = (long long)w * h;
    const int samples = npx < 20000 ? (int)npx : 20000;
    int dmin = an_val(d, 0, 0, w, h, bd, packed, nb), dmax = dmin;
    for (int i = 0; i < samples; ++i) {
        const int idx = (int)((i * 10007) % (int)npx);
        const int v = an_val(d, idx % w, idx / w, w, h, bd, packed, nb);
        if (v < dmin) dmin = v;
        if (v > dmax) dmax = v;
    }
    const int range = dmax - dmin > 1 ? dmax - dmin : 1;
    int step = (int)sqrt((double)w * h / 5000);
    if (step < 1) step = 1;
    step = (step / 2) * 2;
    if (step < 2) step = 2;
    double conf = range * 0.02;
    if (conf < 1.0) conf = 1.0;
    int v1 = 0, v2 = 0;
    for (int y = 0; y < h - 1; y += step)
        for (int x = 0; x < w - 1; x += step) {
            const double tl = an_val(d, x, y, w, h, bd, packed, nb), tr = an_val(d, x + 1, y, w, h, bd, packed, nb);
            const double bl = an_val(d, x, y + 1, w, h, bd, packed, nb), br = an_val(d, x + 1, y + 1, w, h, bd, packed, nb);
            const double d1 = fabs(tl - br), d2 = fabs(tr - bl);
            if (fabs(d1 - d2) < conf) continue;
            const double hi4 = fmax(fmax(tl, tr), fmax(bl, br)), lo4 = fmin(fmin(tl, tr), fmin(bl, br));
            if (hi4 - lo4 > range * 0.15) continue;
            const double hd = fabs(tl - tr) + fabs(bl - br), vd = fabs(tl - bl) + fabs(tr - br);
            const double mn = hd < vd ? hd : vd, mxd = hd < vd ? vd : hd;
            if (mxd / (mn > 1.0 ? mn : 1.0) > 4.0) continue;
            if (d2 < d1) v1++; else v2++;
        }
    int gs = (w < h ? w : h) / 256;
    if (gs < 1) gs = 1;
    gs |= 1;
    double s[4] = {0, 0, 0, 0};
    int c[4] = {0, 0, 0, 0};
    for (int y = 0; y < h; y += gs)
        for (int x = 0; x < w; x += gs) {
            const int k = ((y & 1) == 0 && (x & 1) == 0) ? 0 : ((y & 1) == 1 && (x & 1) == 1) ? 1 : ((y & 1) == 0) ? 2 : 3;
            s[k] += an_val(d, x, y, w, h, bd, packed, nb);
            c[k]++;
        }
    const int margin = v1 > v2 ? v1 - v2 : v2 - v1;
    if (margin <= 2 && v1 + v2 < 10) return 0;
    if (v1 >= v2) {
        const double ee = c[0] ? s[0] / c[0] : 0, oo = c[1] ? s[1] / c[1] : 0;
        const double r = ee > 0 ? oo / ee : 0;
        return r > 1.15 ? 1 : 0; /* BGGR : RGGB */
    }
    const double eo = c[2] ? s[2] / c[2] : 0, oe = c[3] ? s[3] / c[3] : 0;
    const double r = eo > 0 ? oe / eo : 0;
    return r > 1.15 ? 3 : 2; /* GBRG : GRBG */
}
