/* CPU check of the sequence PRISM / RCD use for the CFA normalisation
 * (S0, prism.cpp:58-78: (float)v / (float)max with v an integer sample)
 * in place of the correctly rounded division: with y = 1 / max rounded to
 * nearest (computed once on the host),
 *     q0 = v * y;  q = fma(fma(-max, q0, v), y, q0)
 * must equal (float)v / (float)max for every sample value the decoders can
 * return (0 .. 65535; unclamped 8-bit samples of shallow frames included) at
 * every bit depth 1 .. 16 (max = 2^bd - 1): 1,048,576 cases, all of them.
 *
 * Built with -ffp-contract=off; fmaf is the C library's single-rounding fma.
 * Usage: cfa_norm [plant]   ("plant" drops the correction step, a mistake the
 * check must catch).  Exit status 0 when every case matches. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint32_t asu(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char** argv) {
    const int plant = argc > 1 && strcmp(argv[1], "plant") == 0;
    long bad = 0, n = 0;
    for (int bd = 1; bd <= 16; ++bd) {
        const float m = (float)((1u << bd) - 1u);
        const float y = 1.0f / m;
        for (uint32_t v = 0; v <= 65535u; ++v, ++n) {
            const float a = (float)v;
            const float want = a / m;
            const float q0 = a * y;
            const float got = plant ? q0 : fmaf(fmaf(-m, q0, a), y, q0);
            if (asu(got) != asu(want)) {
                if (bad < 5) printf("bd=%d v=%u: %a != %a\n", bd, v, (double)got, (double)want);
                ++bad;
            }
        }
    }
    printf("%ld cases, %ld mismatches\n", n, bad);
    return bad == 0 ? 0 : 1;
}
