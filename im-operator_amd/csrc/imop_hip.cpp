// imop_hip.cpp -- C ABI of the gfx950 demosaic path (include/imop_hip.h):
// argument checks with the reference's error codes, frame parameter setup,
// stream-ordered PRISM / RCD planes, the host-buffer (synchronous) variants.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "epilogue.hpp"
#include "host_common.hpp"
#include "imop_hip.h"
#include "plain.hpp"

namespace imop_hip {

// Minimum frame side per algorithm, as imop::demosaic_cuda checks it
// (src/algorithms/cuda_dispatch.cpp:77-89).
int min_dim_of(int algo) {
    switch (algo) {
    case kSuperFast: return 2;
    case kHqli: return 6;
    case kMg: return 6;
    case kL7: return 8;
    case kDfpd: return 12;
    case kAhd: return 6;
    case kAmaze: return 6;
    case kRcd: return 10;
    case kPrism: return 10;
    default: return 2;
    }
}

uint64_t bayer_bytes(int w, int h, int bd, int packed) {
    if (packed == 2) return ((uint64_t)w * (uint64_t)h + 1u) / 2u * 3u;  // GigE 10-bit: 3 B per pixel pair
    if (packed) return ((uint64_t)w * (uint64_t)h * (uint64_t)bd + 7u) / 8u;
    return (uint64_t)w * (uint64_t)h * (bd <= 8 ? 1u : 2u);
}

uint64_t rgb_bytes(int w, int h, int bd) { return (uint64_t)w * (uint64_t)h * 3u * (bd > 8 ? 2u : 1u); }

int check_args(const void* in, const void* out, int w, int h, int pattern, int algo, int bd, int packed,
               const float* bl, int bl_mode) {
    if (!in || !out) return kNull;
    if (w <= 0 || h <= 0) return kDims;
    if (bd < 1 || bd > 16) return kBitDepth;
    if (pattern < 0 || pattern > 3) return kPattern;
    if (w < min_dim_of(algo) || h < min_dim_of(algo)) return kSmall;
    if (algo < kSuperFast || algo > kPrism) return kInternal;
    // The reference reads packed streams below 8 bits as one byte per pixel
    // (pixel::get_raw), past the end of the stream; leave those to it.
    if (packed && bd < 8) return kInternal;
    // is_packed: 0 unpacked, 1 the reference's packed layouts, 2 GigE Vision
    // 10-bit packing (bit depth 10 only)
    if (packed < 0 || packed > 2) return kInternal;
    if (packed == 2 && bd != 10) return kBitDepth;
    if (bl_mode < 0 || bl_mode > 2 || (bl_mode != 0 && !bl)) return kInternal;
    return kOk;
}

// Black-level offsets by CFA parity (y & 1) * 2 + (x & 1), each max(0, .).
void bl_by_parity(int pattern, const float* offs, int mode, float q[4]) {
    static const int rr[4] = {0, 1, 0, 1}, rc[4] = {0, 1, 1, 0};
    for (int py = 0; py < 2; ++py)
        for (int px = 0; px < 2; ++px) {
            float o;
            if (mode == 1) o = offs[0];
            else {
                const bool r = py == rr[pattern] && px == rc[pattern];
                const bool b = py != rr[pattern] && px != rc[pattern];
                o = r ? offs[0] : b ? offs[3] : (py == rr[pattern] ? offs[1] : offs[2]);
            }
            q[py * 2 + px] = o > 0.0f ? o : 0.0f;
        }
}

FrameParams make_params(const void* in, uint64_t in_bytes, void* out, int w, int h, int pattern, int bd,
                        int packed, const float* bl, int bl_mode, int algo, int* mode) {
    FrameParams f{};
    f.in = (const uint8_t*)in;
    f.out = (uint8_t*)out;
    // Packed groups are bounded by the stream length pixel::
    // compute_packed_byte_size gives (pixel_utils.hpp:65-66), not by the
    // caller's buffer, which may be larger.
    (void)in_bytes;
    f.in_bytes = bayer_bytes(w, h, bd, packed);
    f.w = w;
    f.h = h;
    f.bd = bd;
    f.pattern = pattern;
    f.out16 = bd > 8;
    f.maxv = (int)((1u << bd) - 1u);
    if (bd <= 8) f.fmt = FMT_U8;
    else if (!packed) f.fmt = FMT_U16;
    else if (packed == 2) f.fmt = FMT_G10;
    else f.fmt = bd == 12 ? FMT_P12 : bd == 10 ? FMT_P10 : FMT_PGEN;
    f.bl_on = bl_mode != 0;
    if (f.bl_on) bl_by_parity(pattern, bl, bl_mode, f.bl_q);
    // With black level the demosaic runs on the unpacked buffer.
    const bool eff_packed = packed && !f.bl_on;
    const bool byte_path = bd <= 8 && !eff_packed;
    f.byte_path = byte_path ? 1 : 0;
    if (algo == kHqli) {
        *mode = MODE_HQLI;
        f.clamp_hi = byte_path ? 255 : f.maxv;
    } else if (algo == kMg) {
        *mode = byte_path ? MODE_MG_INT8 : MODE_MG_EXACT;
        f.clamp_hi = byte_path ? 255 : f.maxv;
    } else {
        *mode = -1;
        f.clamp_hi = f.maxv;
    }
    return f;
}

// PRISM / RCD planes: every call takes its own stream-ordered allocation on
// the calling stream and releases it on that stream after its kernels.
//  * Eager calls allocate from a library-owned pool per device whose release
//    threshold is unlimited, so memory released by one call stays reserved
//    and the next call of the same frame size maps nothing new.  The pool
//    reuses a released block on another stream only once that stream is
//    ordered after the release, so calls on several streams or host threads
//    never share planes while either is running.
//  * Under stream capture the allocation and its release become nodes of the
//    captured graph (hipMallocAsync / hipFreeAsync), so each graph owns its
//    planes and no later eager call can move or free them.
struct PlanePools {
    std::mutex mu;
    std::vector<hipMemPool_t> pool;
};
PlanePools g_planes;

hipError_t plane_pool(hipMemPool_t* out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_planes.mu);
    if ((int)g_planes.pool.size() <= dev) g_planes.pool.resize(dev + 1, nullptr);
    if (!g_planes.pool[dev]) {
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t p = nullptr;
        e = hipMemPoolCreate(&p, &props);
        if (e != hipSuccess) return e;
        uint64_t keep = UINT64_MAX;
        e = hipMemPoolSetAttribute(p, hipMemPoolAttrReleaseThreshold, &keep);
        if (e != hipSuccess) {
            (void)hipMemPoolDestroy(p);
            return e;
        }
        g_planes.pool[dev] = p;
    }
    *out = g_planes.pool[dev];
    return hipSuccess;
}

hipError_t planes_alloc(size_t bytes, hipStream_t st, void** out) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hipError_t e = hipStreamIsCapturing(st, &cs);
    if (e != hipSuccess) return e;
    if (cs != hipStreamCaptureStatusNone) return hipMallocAsync(out, bytes, st);
    hipMemPool_t pool = nullptr;
    e = plane_pool(&pool);
    if (e != hipSuccess) return e;
    return hipMallocFromPoolAsync(out, bytes, pool, st);
}

int launch_frame(const FrameParams& f, int algo, int mode, void* prism_ws, hipStream_t st, const Epilogue* epi) {
    hipError_t he;
    const bool with_epi = epi && epi->flags;
    // fused into the register-resident kernel, which keeps the part of the
    // gamma table its outputs can index in LDS (u8 RGB: 256 entries; u16 RGB:
    // max + 1 <= 4096, bit depths up to 12); otherwise the epilogue runs as
    // its own pass
    const uint32_t lds_entries = f.out16 ? (uint32_t)f.maxv + 1u : 256u;
    if (with_epi && (algo == kHqli || algo == kMg) && fast_path_ok(f) && (!(epi->flags & EPI_GAMMA) || lds_entries <= 4096u)) {
        he = launch_fast(f, mode, st, epi);
        return he == hipSuccess ? kOk : kInternal;
    }
    if (algo == kSuperFast) {
        he = launch_super_fast(f, st);
    } else if (algo == kL7) {
        he = launch_l7(f, st);
    } else if (algo == kDfpd) {
        he = launch_dfpd(f, st);
    } else if (algo == kAhd) {
        he = launch_ahd(f, st);
    } else if (algo == kAmaze) {
        he = launch_amaze(f, st);
    } else if (algo == kPrism || algo == kRcd) {
        // the PRISM / RCD passes index their planes with 32-bit offsets
        if ((uint64_t)f.w * (uint64_t)f.h >= (1ull << 31)) return kInternal;
        void* ws = prism_ws;
        if (!ws) {
            he = planes_alloc(prism_workspace_bytes(f.w, f.h), st, &ws);
            if (he != hipSuccess) return kInternal;
        }
        he = launch_prism(f, ws, st, algo == kRcd);
        if (!prism_ws) {
            const hipError_t fe = hipFreeAsync(ws, st);
            if (he == hipSuccess) he = fe;
        }
    } else if (fast_path_ok(f)) {
        he = launch_fast(f, mode, st);
    } else {
        he = launch_plain(f, mode, st);
    }
    if (he == hipSuccess && with_epi) he = launch_rgb_epilogue(f.out, (uint64_t)f.w * (uint64_t)f.h, f.out16, *epi, st);
    return he == hipSuccess ? kOk : kInternal;
}

int run_demosaic(const void* d_in, uint64_t in_bytes, void* d_out, uint64_t out_bytes, int w, int h, int pattern,
                 int algo, int bd, int packed, const float* bl, int bl_mode, hipStream_t st, const Epilogue* epi) {
    int e = check_args(d_in, d_out, w, h, pattern, algo, bd, packed, bl, bl_mode);
    if (e != kOk) return e;
    if (in_bytes < bayer_bytes(w, h, bd, packed) || out_bytes < rgb_bytes(w, h, bd)) return kInternal;
    // u16 samples and u16 RGB are accessed as naturally aligned 16-bit words
    if (bd > 8 && ((!packed && ((uintptr_t)d_in & 1)) || ((uintptr_t)d_out & 1))) return kInternal;
    int mode = -1;
    FrameParams f = make_params(d_in, in_bytes, d_out, w, h, pattern, bd, packed, bl, bl_mode, algo, &mode);
    return launch_frame(f, algo, mode, nullptr, st, epi);
}

namespace {

// Per-host-thread device buffers for the host-pointer entry points, like the
// reference's thread_local CudaWorkspace (cuda_dispatch.cpp:51).
struct HostStaging {
    void* d_in = nullptr;
    size_t in_cap = 0;
    void* d_out = nullptr;
    size_t out_cap = 0;
    int dev = -1;
    ~HostStaging() {
        if (d_in) (void)hipFree(d_in);
        if (d_out) (void)hipFree(d_out);
    }
    bool ensure(size_t in_b, size_t out_b) {
        int dev_now = 0;
        if (hipGetDevice(&dev_now) != hipSuccess) return false;
        if (dev_now != dev) {
            if (d_in) (void)hipFree(d_in);
            if (d_out) (void)hipFree(d_out);
            d_in = d_out = nullptr;
            in_cap = out_cap = 0;
            dev = dev_now;
        }
        if (in_cap < in_b) {
            if (d_in) (void)hipFree(d_in);
            d_in = nullptr;
            in_cap = 0;
            if (hipMalloc(&d_in, in_b) != hipSuccess) return false;
            in_cap = in_b;
        }
        if (out_cap < out_b) {
            if (d_out) (void)hipFree(d_out);
            d_out = nullptr;
            out_cap = 0;
            if (hipMalloc(&d_out, out_b) != hipSuccess) return false;
            out_cap = out_b;
        }
        return true;
    }
};
thread_local HostStaging t_staging;

}  // namespace

int device_count_cached() {
    static std::once_flag once;
    static int count = 0;
    std::call_once(once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) == hipSuccess && n > 0) count = n;
    });
    return count;
}

}  // namespace imop_hip

using namespace imop_hip;

extern "C" {

int imop_hip_device_count(void) { return device_count_cached(); }

const char* imop_hip_device_name(void) {
    static char name[256] = "N/A";
    static std::once_flag once;
    if (device_count_cached() == 0) return "N/A";
    std::call_once(once, [] {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, 0) == hipSuccess) {
            std::strncpy(name, p.name, sizeof(name) - 1);
            name[sizeof(name) - 1] = '\0';
        } else {
            std::strncpy(name, "Unknown", sizeof(name) - 1);
        }
    });
    return name;
}

int imop_hip_synchronize(void) {
    if (device_count_cached() == 0) return kInternal;
    return hipDeviceSynchronize() == hipSuccess ? kOk : kInternal;
}

int imop_hip_demosaic(const void* d_in, uint64_t in_bytes, void* d_out, uint64_t out_bytes, int width, int height,
                      int pattern, int algorithm, int bit_depth, int is_packed, const float* black_level,
                      int black_level_mode, void* stream) {
    if (device_count_cached() == 0) return kInternal;
    return run_demosaic(d_in, in_bytes, d_out, out_bytes, width, height, pattern, algorithm, bit_depth, is_packed,
                        black_level, black_level_mode, (hipStream_t)stream);
}

int imop_hip_demosaic_host(const uint8_t* bayer, uint8_t* rgb, int width, int height, int pattern, int algorithm,
                           int bit_depth, int is_packed, const float* black_level, int black_level_mode) {
    if (device_count_cached() == 0) return kInternal;
    int e = check_args(bayer, rgb, width, height, pattern, algorithm, bit_depth, is_packed, black_level,
                       black_level_mode);
    if (e != kOk) return e;
    const uint64_t ib = bayer_bytes(width, height, bit_depth, is_packed);
    const uint64_t ob = rgb_bytes(width, height, bit_depth);
    if (!t_staging.ensure(ib, ob)) return kInternal;
    if (hipMemcpy(t_staging.d_in, bayer, ib, hipMemcpyHostToDevice) != hipSuccess) return kInternal;
    e = run_demosaic(t_staging.d_in, ib, t_staging.d_out, ob, width, height, pattern, algorithm, bit_depth,
                     is_packed, black_level, black_level_mode, nullptr);
    if (e != kOk) return e;
    if (hipMemcpy(rgb, t_staging.d_out, ob, hipMemcpyDeviceToHost) != hipSuccess) return kInternal;
    return kOk;
}

int imop_hip_set_launch_shape(int rows_per_wave, int waves_per_block, int xcd_remap) {
    set_launch_shape(rows_per_wave, waves_per_block, xcd_remap);
    return kOk;
}

int imop_hip_get_launch_shape(int* rows_per_wave, int* waves_per_block, int* xcd_remap) {
    get_launch_shape(rows_per_wave, waves_per_block, xcd_remap);
    return kOk;
}

int imop_hip_black_level(void* d_data, int width, int height, int pattern, int algorithm, int bit_depth,
                         const float* offsets, void* stream) {
    // black_level::BlackLevelError: NullInput 1, InvalidDimensions 2,
    // InvalidBitDepth 3, InternalError 6 (include/black_level/types.hpp).
    if (device_count_cached() == 0) return kInternal;
    if (!d_data || !offsets) return kNull;
    if (width <= 0 || height <= 0) return kDims;
    if (bit_depth < 1 || bit_depth > 16) return kBitDepth;
    if (pattern < 0 || pattern > 3 || algorithm < 0 || algorithm > 1) return kInternal;
    float q[4];
    bl_by_parity(pattern, offsets, algorithm == 1 ? 1 : 2, q);
    return launch_black_level((uint8_t*)d_data, width, height, bit_depth, q, (hipStream_t)stream) == hipSuccess
               ? kOk
               : kInternal;
}

int imop_hip_black_level_host(uint8_t* data, int width, int height, int pattern, int algorithm, int bit_depth,
                              const float* offsets) {
    if (device_count_cached() == 0) return kInternal;
    if (!data || !offsets) return kNull;
    if (width <= 0 || height <= 0) return kDims;
    if (bit_depth < 1 || bit_depth > 16) return kBitDepth;
    const uint64_t nb = bayer_bytes(width, height, bit_depth, 0);
    if (!t_staging.ensure(nb, 1)) return kInternal;
    if (hipMemcpy(t_staging.d_in, data, nb, hipMemcpyHostToDevice) != hipSuccess) return kInternal;
    int e = imop_hip_black_level(t_staging.d_in, width, height, pattern, algorithm, bit_depth, offsets, nullptr);
    if (e != kOk) return e;
    if (hipMemcpy(data, t_staging.d_in, nb, hipMemcpyDeviceToHost) != hipSuccess) return kInternal;
    return kOk;
}

}  // extern "C"
