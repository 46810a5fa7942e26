// Host check of the batched kernels' frame-major placement
// (frame_major_local in batch.hpp).  For every frame item count n =
// 1..20000 and a spread of larger ones: the slots 0 .. ceil(n / 8) - 1 of
// the 8 XCDs hand out every item 0 .. n-1 exactly once, with
// 8 * ceil(n / 8) - n idle slots; XCD x's items are one contiguous run,
// after XCD x-1's, walked forward on XCDs 0-3 and backward on 4-7; and slot
// j of XCD x is the item the single-frame kernel's placement gives its
// workgroup 8j + x in a launch of n workgroups.  `plant` gives every XCD q
// slots (drops the remainder items), which must be caught.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "batch.hpp"

using imop_hip::frame_major_local;

static bool plant = false;

static bool local_of(uint32_t n, uint32_t xcd, uint32_t j, uint32_t& local) {
    if (!plant) return frame_major_local(n, xcd, j, local);
    const uint32_t q = n >> 3;
    if (j >= q) return false;
    local = xcd * q + (xcd >= 4 ? q - 1u - j : j);
    return true;
}

// The single-frame kernel's workgroup placement over a grid of n workgroups
// (block b on XCD b & 7, slot b >> 3), written out independently.
static uint32_t single_frame_item(uint32_t n, uint32_t b) {
    const uint32_t xcd = b & 7u, slot = b >> 3;
    uint32_t start = 0;
    for (uint32_t x = 0; x < xcd; ++x) start += n / 8 + (x < n % 8 ? 1u : 0u);
    const uint32_t cnt = n / 8 + (xcd < n % 8 ? 1u : 0u);
    return start + (xcd >= 4 ? cnt - 1u - slot : slot);
}

static uint64_t frames = 0, bad = 0;

static void check(uint32_t n) {
    ++frames;
    const uint32_t slots = (n + 7u) / 8u;
    std::vector<uint8_t> seen(n, 0);
    uint32_t idle = 0, prev_hi = 0;
    bool ok = true;
    for (uint32_t x = 0; x < 8 && ok; ++x) {
        uint32_t lo = UINT32_MAX, hi = 0, cnt = 0, last = 0;
        for (uint32_t j = 0; j < slots && ok; ++j) {
            uint32_t local = 0;
            if (!local_of(n, x, j, local)) { ++idle; continue; }
            if (local >= n || seen[local]) { ok = false; break; }
            seen[local] = 1;
            if (8u * j + x < n && local != single_frame_item(n, 8u * j + x)) ok = false;
            if (cnt > 0 && local != (x >= 4 ? last - 1u : last + 1u)) ok = false;  // walk direction
            last = local;
            lo = local < lo ? local : lo;
            hi = local > hi ? local : hi;
            ++cnt;
        }
        if (cnt > 0) {
            if (hi - lo + 1u != cnt || lo != prev_hi) ok = false;  // one contiguous run after XCD x-1's
            prev_hi = hi + 1u;
        }
    }
    for (uint32_t i = 0; i < n && ok; ++i) ok = seen[i] != 0;
    if (ok && idle != 8u * slots - n) ok = false;
    if (!ok && bad++ < 5) printf("mismatch: n=%u\n", n);
}

int main(int argc, char** argv) {
    plant = argc > 1 && strcmp(argv[1], "plant") == 0;
    for (uint32_t n = 1; n <= 20000; ++n) check(n);
    for (uint32_t n = 20001; n < 4000000; n = n * 3 / 2 + 7) check(n);
    printf("%llu frames, %llu mismatches\n", (unsigned long long)frames, (unsigned long long)bad);
    return bad == 0 ? 0 : 1;
}
