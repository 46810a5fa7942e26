// median9.hpp -- the 3x3 exchange network of prism_detail::median_3x3
// (src/algorithms/prism.cpp:9-21), which apply_median_filter_3x3
// (src/algorithms/ahd.cpp:7-40) repeats exchange for exchange.
#pragma once

#include <hip/hip_runtime.h>

namespace imop_hip {

// prism_detail::median_3x3 (prism.cpp:9-21): the same exchange network.  (It
// is not a true median network -- it fails the 0-1 principle -- so it is kept
// exactly.)  Each exchange "if (a > b) swap" is min / max: with no NaNs the
// two differ only in which sign a +-0 tie keeps, and a signed zero selected
// here leaves g + median, the clamp and the rounding unchanged.
__device__ __forceinline__ float median9(float v[9]) {
#define IMOP_CX(a, b)                        \
    {                                        \
        const float ta = v[a], tb = v[b];    \
        v[a] = fminf(ta, tb);                \
        v[b] = fmaxf(ta, tb);                \
    }
    IMOP_CX(0, 1) IMOP_CX(3, 4) IMOP_CX(6, 7) IMOP_CX(1, 2) IMOP_CX(4, 5) IMOP_CX(7, 8)
    IMOP_CX(0, 1) IMOP_CX(3, 4) IMOP_CX(6, 7) IMOP_CX(0, 3) IMOP_CX(4, 7) IMOP_CX(1, 4)
    IMOP_CX(3, 6) IMOP_CX(2, 5) IMOP_CX(5, 8) IMOP_CX(1, 3) IMOP_CX(4, 6) IMOP_CX(2, 5)
    IMOP_CX(2, 3) IMOP_CX(5, 6) IMOP_CX(2, 4) IMOP_CX(3, 5) IMOP_CX(3, 4)
#undef IMOP_CX
    return v[4];
}

// The network's first nine exchanges sort each row triple (v[0..2], v[3..5],
// v[6..8]) on its own, by (0,1) (1,2) (0,1); they touch disjoint triples, so
// a sorted triple can be formed once and shared by the two pixels whose
// windows hold that row.  sort3 is that triple sort; median9_of_sorted_rows
// runs the remaining fourteen exchanges in the network's order, so
// median9_of_sorted_rows(rows sorted by sort3) == median9(rows) bit for bit.
__device__ __forceinline__ void sort3(float& a, float& b, float& c) {
    float t = fminf(a, b);
    b = fmaxf(a, b);
    a = t;
    t = fminf(b, c);
    c = fmaxf(b, c);
    b = t;
    t = fminf(a, b);
    b = fmaxf(a, b);
    a = t;
}

__device__ __forceinline__ float median9_of_sorted_rows(float v[9]) {
#define IMOP_CX(a, b)                        \
    {                                        \
        const float ta = v[a], tb = v[b];    \
        v[a] = fminf(ta, tb);                \
        v[b] = fmaxf(ta, tb);                \
    }
    IMOP_CX(0, 3) IMOP_CX(4, 7) IMOP_CX(1, 4) IMOP_CX(3, 6) IMOP_CX(2, 5) IMOP_CX(5, 8) IMOP_CX(1, 3)
    IMOP_CX(4, 6) IMOP_CX(2, 5) IMOP_CX(2, 3) IMOP_CX(5, 6) IMOP_CX(2, 4) IMOP_CX(3, 5) IMOP_CX(3, 4)
#undef IMOP_CX
    return v[4];
}

}  // namespace imop_hip
