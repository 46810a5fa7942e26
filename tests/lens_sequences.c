/* CPU check of the two float sequences the register-resident lens-shading
 * decode uses in place of the correctly rounded sqrt and division
 * (sqrt_rn / div_rn in im-operator_amd/csrc/fast_impl.hpp), over the range
 * the host admits for them (make_lens_params in isp_kernels.hip: centre
 * coordinates 0 or of magnitude in [2^-20, 2^40], so sums of squares t are 0
 * or in [2^-86, 2^82], roots s in [2^-43, 2^41] and max_r in [2^-20, 2^41]).
 *
 *   sqrt_rn: the hardware estimate is within 1 ulp of sqrt(t); for each t
 *            the correction is run from all three such estimates (the
 *            correctly rounded root and both neighbours) and must return the
 *            correctly rounded root.
 *   div_rn:  from y = 1/m rounded to nearest, two fused remainder steps on
 *            s * y must give s / m rounded to nearest.
 *
 * Built with -ffp-contract=off; fmaf is the C library's single-rounding fma.
 * Usage: lens_sequences [stride] [plant]   (every stride-th float of each
 * binade; exit status 0 when every case matches; "plant" returns the
 * uncorrected quotient s * y, a mistake the check must catch). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float asf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t asu(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float sqrt_fix(float t, float s) {
    const float sm = asf(asu(s) - 1u), sp = asf(asu(s) + 1u);
    const float lo = fmaf(-sm, s, t) <= 0.0f ? sm : s;
    return fmaf(-sp, s, t) > 0.0f ? sp : lo;
}

static int g_plant = 0; /* planted mistake: the uncorrected product s * y */

static float div_seq(float a, float b, float y) {
    const float q0 = a * y;
    if (g_plant) return q0;
    const float q1 = fmaf(fmaf(-b, q0, a), y, q0);
    return fmaf(fmaf(-b, q1, a), y, q1);
}

/* the host's reciprocal: the double quotient's float or a nearer neighbour */
static float recip_rn(float m) {
    const float y = (float)(1.0 / (double)m);
    const float c[3] = {nextafterf(y, 0.0f), y, nextafterf(y, INFINITY)};
    int best = 1;
    double err = fabs((double)m * (double)c[1] - 1.0);
    for (int k = 0; k < 3; k += 2) {
        const double e = fabs((double)m * (double)c[k] - 1.0);
        if (e < err) { err = e; best = k; }
    }
    return c[best];
}

int main(int argc, char** argv) {
    const uint32_t stride = argc > 1 ? (uint32_t)atoi(argv[1]) : 1001u;
    g_plant = argc > 2 && strcmp(argv[2], "plant") == 0;
    long n = 0, bad = 0;
    /* sqrt: t = 0 and every binade of [2^-86, 2^82] */
    for (int d = -1; d <= 1; ++d)
        if (asu(sqrt_fix(0.0f, asf(asu(0.0f) + (d > 0 ? 1u : 0u)))) != 0u) { ++bad; printf("sqrt(0) from %d\n", d); }
    for (int e = -86; e <= 82; ++e) {
        const uint32_t base = asu(ldexpf(1.0f, e));
        for (uint32_t k = (uint32_t)e & 7u; k < (1u << 23); k += stride) {
            const float t = asf(base + k), want = sqrtf(t);
            for (int d = -1; d <= 1; ++d) {
                ++n;
                if (asu(sqrt_fix(t, asf(asu(want) + (uint32_t)d))) != asu(want)) {
                    if (bad++ < 5) printf("sqrt t=%.9g estimate %+d ulp\n", t, d);
                }
            }
        }
    }
    /* division: divisors at the ends of the range, all-ones significands,
     * powers of two, and frame-like radii; dividends over [2^-43, 2^41] */
    float ms[48];
    int nm = 0;
    const float fixed[] = {0x1p-20f, 0x1p41f, 1.0f, 2.0f, 3.0f, 0x1.fffffep0f, 0x1.fffffep23f, 0x1.fffffep-20f,
                           0x1.fffffep40f, 2528.1003f, 2529.0f, 5050.3213f, 0.70710678f, 1e-3f, 123456.78f, 4096.0f};
    for (unsigned i = 0; i < sizeof fixed / sizeof fixed[0]; ++i) ms[nm++] = fixed[i];
    srand(7);
    while (nm < 48) {
        const float cx = (float)(rand() % 8192) * ((rand() & 1) ? 0.5f : 0.37f), cy = (float)(rand() % 6144) * 0.5f;
        const float m = sqrtf(fmaf(cx, cx, cy * cy));
        if (m > 0.0f) ms[nm++] = m;
    }
    for (int i = 0; i < nm; ++i) {
        const float m = ms[i], y = recip_rn(m);
        for (int e = -43; e <= 40; ++e) {
            const uint32_t base = asu(ldexpf(1.0f, e));
            for (uint32_t k = (uint32_t)(e + i) & 15u; k < (1u << 23); k += stride) {
                const float s = asf(base + k);
                ++n;
                if (asu(div_seq(s, m, y)) != asu(s / m)) {
                    if (bad++ < 5) printf("div s=%.9g m=%.9g\n", s, m);
                }
            }
        }
    }
    printf("%ld cases, %ld mismatches\n", n, bad);
    return bad != 0;
}
