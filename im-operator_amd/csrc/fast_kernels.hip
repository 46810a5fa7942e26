// fast_kernels.hip -- host-side selection and parameters of the
// register-resident HQLI / MG kernel (fast_impl.hpp, instantiated per raw
// layout in fast_u8 / fast_u16 / fast_p12 / fast_p10.hip).
#include <atomic>

#include "device_common.hpp"
#include "plain.hpp"

namespace imop_hip {

// Launch shape (set by the autotuner through imop_hip_set_launch_shape), one
// word so that a launch never sees half of a concurrent change:
// rows | wpb << 16 | xcd_remap << 24.  rows: rows per wave (2 or 4; 0 =
// automatic); wpb: waves per group (4; groups of 2 waves measured no faster
// and are not instantiated).  xcd_remap bit 0:
// XCD-contiguous group order; bit 1: strip-fastest group order.  Automatic
// rows: 2.  Groups of 4 waves x 2 rows (8 output rows) decode 14 raw rows
// where 4 x 4 groups decode 22 for 16, but the kernel is bound by its RGB
// stores, and 8-row groups keep fewer rows of the frame in flight at once:
// 12 MP MG frame (RAW12 -> u16) 16.2-17.1 us against 17.1-17.9 us for 4-row
// waves on the same boxes; small frames gain the same way from twice the
// waves.
static std::atomic<uint32_t> g_shape{0u | (4u << 16) | (3u << 24)};

void set_launch_shape(int rows, int wpb, int xcd_remap) {
    const bool rows_ok = rows == 0 || rows == 2 || rows == 4;
    uint32_t cur = g_shape.load();
    for (;;) {
        int r = (int)(cur & 0xFFFFu), w = (int)((cur >> 16) & 0xFFu), x = (int)(cur >> 24);
        if (rows_ok) r = rows;
        if (wpb == 4) w = wpb;
        if (xcd_remap >= 0 && xcd_remap <= 3) x = xcd_remap;
        const uint32_t next = (uint32_t)r | ((uint32_t)w << 16) | ((uint32_t)x << 24);
        if (g_shape.compare_exchange_weak(cur, next)) return;
    }
}

void get_launch_shape(int* rows, int* wpb, int* xcd_remap) {
    const uint32_t s = g_shape.load();
    if (rows) *rows = (int)(s & 0xFFFFu);
    if (wpb) *wpb = (int)((s >> 16) & 0xFFu);
    if (xcd_remap) *xcd_remap = (int)(s >> 24);
}

hipError_t launch_fast_u8(const FastParams& f, int mode, hipStream_t st);
hipError_t launch_fast_u16(const FastParams& f, int mode, hipStream_t st);
hipError_t launch_fast_p12(const FastParams& f, int mode, hipStream_t st);
hipError_t launch_fast_g10(const FastParams& f, int mode, hipStream_t st);
hipError_t launch_fast_p10(const FastParams& f, int mode, hipStream_t st);

bool fast_path_ok(const FrameParams& f) {
    if ((f.w % 4 != 0 && f.w % 4 != 2) || f.w < 8 || f.h < 6) return false;
    if (((uintptr_t)f.in | (uintptr_t)f.out) & 7) return false;
    if ((uint64_t)f.w * f.h * 6u >= (1ull << 31) || f.in_bytes >= (1ull << 31)) return false;  // 32-bit offsets
    if (f.fmt != FMT_U8 && f.fmt != FMT_U16 && f.fmt != FMT_P12 && f.fmt != FMT_P10 && f.fmt != FMT_G10) return false;
    if (f.bl_on) {
        for (int k = 0; k < 4; ++k) {
            const float o = f.bl_q[k];
            if (!(o >= 0.0f && o <= 65535.0f) || o != (float)(int)o) return false;
        }
    }
    return true;
}

// Groups of wpb waves x th rows over nsx strips.
FastParams make_fast(const FrameParams& f, int th, int wpb, int xcd) {
    FastParams q{};
    q.in = f.in;
    q.out = f.out;
    q.in_bytes = f.in_bytes;
    q.w = f.w;
    q.h = f.h;
    q.bd = f.bd;
    q.maxv = f.maxv;
    q.clamp_hi = f.clamp_hi;
    q.pattern = f.pattern;
    q.out16 = f.out16;
    q.bl_on = f.bl_on;
    for (int k = 0; k < 4; ++k) q.bl_i[k] = f.bl_on ? (int)f.bl_q[k] : 0;
    q.th = th == 2 ? 2 : 4;
    q.wpb = 4;
    (void)wpb;
    q.xcd_remap = xcd;
    q.aligned = 1;
    q.nsx = strip_count(f.w, f.out16 != 0);
    q.nby = (f.h + q.th * q.wpb - 1) / (q.th * q.wpb);  // group rows
    q.nwaves = q.nsx * q.nby * q.wpb;
    // group placement divides the block index by the strip or group-row count
    // (exact while blocks x divisor < 2^40; else 0: the kernel divides)
    const uint64_t d = (xcd & 2) ? (uint64_t)q.nsx : (uint64_t)q.nby;
    const uint64_t blocks = (uint64_t)q.nsx * (uint64_t)q.nby;
    q.place_mul = place_reciprocal(blocks, d);
    return q;
}

hipError_t launch_fast(const FrameParams& f, int mode, hipStream_t st, const Epilogue* epi, const LensParams* lens) {
    int rows, wpb, xcd;
    get_launch_shape(&rows, &wpb, &xcd);
    const bool with_epi = (epi && epi->flags) || (lens && lens->on);
    if (with_epi) {  // the epilogue form: 4-row waves in groups of 4
        rows = 4;
        wpb = 4;
    } else if (rows == 0) {
        rows = 2;
    }
    FastParams q = make_fast(f, rows, wpb, xcd);
    if (epi && epi->flags) q.epi = *epi;
    if (lens && lens->on) q.lens = *lens;
    switch (f.fmt) {
    case FMT_U8: return launch_fast_u8(q, mode, st);
    case FMT_U16: return launch_fast_u16(q, mode, st);
    case FMT_P12: return launch_fast_p12(q, mode, st);
    case FMT_G10: return launch_fast_g10(q, mode, st);
    default: return launch_fast_p10(q, mode, st);
    }
}

}  // namespace imop_hip
