// Host check of the work-item placement shared by the single-frame and the
// batched kernels (place_reciprocal / place_split in device_common.hpp).  For
// every divisor d = 1..65535: the largest item count n the reciprocal
// accepts covers every count the kernels form with that divisor (at most
// 65535 x d items, or the 2^40 bound), and at that n the items where a
// truncation or overflow error would show first (the last item, the last
// ones of the final full quotient, the first of the next) and spread items
// come out as the exact quotient and remainder in both orders; the plain
// division (m = 0) as well.  `plant` uses floor(2^40 / d) instead of the
// ceiling, which must be caught.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "device_common.hpp"

using imop_hip::place_reciprocal;
using imop_hip::place_split;

static uint64_t cases = 0, bad = 0;

static void check(uint32_t i, uint32_t d, uint32_t other, uint64_t m) {
    for (int order = 0; order < 2; ++order) {
        const bool sf = order == 1;
        int sx = -1, g = -1;
        place_split(i, sf, sf ? d : other, sf ? other : d, m, sx, g);
        const uint32_t q = i / d, r = i % d;
        const bool ok = sf ? ((uint32_t)g == q && (uint32_t)sx == r) : ((uint32_t)sx == q && (uint32_t)g == r);
        ++cases;
        if (!ok && bad++ < 5) printf("mismatch: i=%u d=%u order=%d -> sx=%d g=%d\n", i, d, order, sx, g);
    }
}

int main(int argc, char** argv) {
    const bool plant = argc > 1 && strcmp(argv[1], "plant") == 0;
    uint32_t seed = 12345u;
    for (uint32_t d = 1; d <= 65535u; ++d) {
        // largest accepted count (acceptance is monotone in n), up to 2^31 items
        uint64_t lo = 0, hi = 1ull << 31;
        while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) / 2;
            if (place_reciprocal(mid, d) != 0) lo = mid;
            else hi = mid - 1;
        }
        const uint64_t n = lo;
        uint64_t need = 65535ull * d;
        if (need > ((1ull << 40) - 1) / d) need = ((1ull << 40) - 1) / d;
        if (n < need || n == 0) {
            printf("reciprocal refused a count the kernels form: d=%u accepted up to %llu, need %llu\n", d,
                   (unsigned long long)n, (unsigned long long)need);
            return 1;
        }
        uint64_t m = place_reciprocal(n, d);
        if (plant) m = (1ull << 40) / d;
        const uint64_t top = ((n - 1) / d) * d;  // first item of the last (possibly partial) quotient
        const uint64_t pts[] = {n - 1, top ? top - 1 : 0, top, top > d ? top - d : 0, top > d ? top - d + 1 : 0, 0,
                                d - 1 < n ? d - 1 : 0};
        for (uint64_t p : pts) check((uint32_t)p, d, 7u, m);
        for (int k = 0; k < 8; ++k) {
            seed = seed * 1664525u + 1013904223u;
            check((uint32_t)((uint64_t)seed % n), d, 65535u, m);
        }
        check(5u % d, d, 3u, 0);  // m = 0: the plain division
    }
    printf("%llu cases, %llu mismatches\n", (unsigned long long)cases, (unsigned long long)bad);
    return bad == 0 ? 0 : 1;
}
