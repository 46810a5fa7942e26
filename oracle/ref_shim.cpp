// ref_shim.cpp -- C entry points over the reference's own CPU demosaic and
// black-level functions, so tests can call the compiled reference through
// ctypes.  TEST INFRASTRUCTURE ONLY: built by oracle/build_ref.sh together
// with the reference sources under /root/reference into oracle/_ref/, used to
// pin the restatement in imop_oracle.c and as the timed CPU baseline.
#include <cstddef>
#include <cstdint>

#include "imop/algorithms.hpp"
#include "imop/pixel_utils.hpp"
#include "black_level/algorithms.hpp"
#include "common.hpp"  // src/algorithms/common.hpp: detail::thread_local_pool

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// imop::demosaic_cpu, src/algorithms/dispatch.cpp:115-128.
//
// process_prism takes `auto& plane = pool.get(total, k)` for k = 0..10 in
// turn (prism.cpp:44-55); FloatBufferPool::get grows its outer vector one
// slot at a time (common.hpp:16-23), which moves the earlier slots and
// leaves the earlier references dangling, so the first PRISM call on a
// thread writes through a dangling reference and crashes.  Growing the
// pool to 16 slots before the call keeps every reference valid; the planes
// themselves are still sized and zero-filled by process_prism as before.
int ref_demosaic_cpu(const uint8_t* in, uint8_t* rgb, int w, int h, int pattern, int algo, int bd,
                     int packed) {
    imop::detail::thread_local_pool().get(1, 15);
    return static_cast<int>(imop::demosaic_cpu(in, rgb, w, h, static_cast<imop::BayerPattern>(pattern),
                                               static_cast<imop::DemosaicAlgorithm>(algo), bd, packed != 0));
}

// black_level::process_global / process_per_channel,
// src/black_level/global.cpp:6-24, src/black_level/per_channel.cpp:6-33
int ref_black_level(uint8_t* data, int w, int h, int pattern, int bd, const float* offs, int per_channel) {
    black_level::BlackLevelParams p;
    p.r_offset = offs[0];
    p.gr_offset = offs[1];
    p.gb_offset = offs[2];
    p.b_offset = offs[3];
    auto pat = static_cast<imop::BayerPattern>(pattern);
    auto e = per_channel ? black_level::process_per_channel(data, w, h, pat, bd, p)
                         : black_level::process_global(data, w, h, pat, bd, p);
    return static_cast<int>(e);
}

// pixel::get_packed_raw, include/imop/pixel_utils.hpp:58-118
int ref_get_packed_raw(const uint8_t* data, int x, int y, int w, int bd, size_t nbytes, int h) {
    return imop::pixel::get_packed_raw(data, x, y, w, bd, nbytes, h);
}

// pixel::get_raw, include/imop/pixel_utils.hpp:139-150
int ref_get_raw(const uint8_t* data, int x, int y, int w, int bd, int packed, size_t nbytes, int h) {
    return imop::pixel::get_raw(data, x, y, w, bd, packed != 0, nbytes, h);
}

// The caller-side unpack that precedes black level: get_raw for every pixel
// into a u8 (bd <= 8) or u16 buffer.
void ref_unpack(const uint8_t* data, int w, int h, int bd, int packed, uint8_t* out) {
    const size_t n = static_cast<size_t>(w) * h;
    for (size_t i = 0; i < n; ++i) {
        int v = imop::pixel::get_raw(data, static_cast<int>(i % w), static_cast<int>(i / w), w, bd, packed != 0, 0, h);
        if (bd <= 8) out[i] = static_cast<uint8_t>(v);
        else imop::pixel::write_u16(out, 2 * i, static_cast<uint16_t>(v));
    }
}

int ref_has_avx2() { return imop::has_avx2() ? 1 : 0; }

int ref_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ref_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

}  // extern "C"
