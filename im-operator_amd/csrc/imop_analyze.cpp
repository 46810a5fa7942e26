// imop_analyze.cpp -- C ABI of the raw-data analyzer (include/imop_hip.h):
// imop::analyze_data and imop::guess_pattern (src/analyzer.cpp:13-321) on a
// buffer in device memory.  The GPU gathers the statistics the reference's
// heuristics look at (analyze_kernels.hip); the decisions below are the
// reference's, taken on those statistics.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "analyze.hpp"
#include "host_common.hpp"
#include "imop_hip.h"

namespace imop_hip {
namespace {

// match_common_resolution (analyzer.cpp:13-25)
size_t common_resolution(size_t px) {
    static const size_t sizes[] = {640 * 480,   720 * 480,   720 * 576,   800 * 600,   1024 * 768,  1280 * 720,
                                   1280 * 800,  1280 * 1024, 1440 * 900,  1920 * 1080, 1920 * 1200, 2048 * 1536,
                                   2560 * 1440, 2560 * 1600, 3840 * 2160, 4096 * 2160, 4096 * 3072, 5120 * 2880,
                                   5320 * 4600, 7680 * 4320, 1006 * 758,  4024 * 3032};
    for (size_t s : sizes)
        if (px == s) return s;
    return 0;
}

// suggest_dimensions (analyzer.cpp:119-149)
std::vector<std::pair<int, int>> suggest_dimensions(size_t px) {
    std::vector<std::pair<int, int>> r;
    if (px == 0) return r;
    static const int cw[] = {640, 720, 800, 1024, 1280, 1440, 1920, 2048, 2560, 3840, 4096, 5120, 7680};
    static const int chs[] = {480, 576, 600, 720, 768, 800, 1024, 1080, 1200, 1440, 1536, 2048, 2160, 3072, 4320};
    for (int w : cw)
        if (px % (size_t)w == 0) {
            const int h = (int)(px / (size_t)w);
            for (int c : chs)
                if (h == c) {
                    r.push_back({w, h});
                    break;
                }
        }
    if (r.empty()) {
        const int s = (int)std::sqrt((double)px);
        for (int w = s - 100; w <= s + 100; ++w)
            if (w > 0 && px % (size_t)w == 0) {
                const int h = (int)(px / (size_t)w);
                if (w >= 64 && h >= 64) {
                    r.push_back({w, h});
                    if (r.size() >= 5) break;
                }
            }
    }
    return r;
}

// Per-device partial-statistics rows and the pinned host AnalyzeStats the
// last kernel writes through its device mapping; one call at a time per device.
struct AnalyzeState {
    std::mutex mu;
    AnalyzeRow* rows = nullptr;         // kAnalyzeRowCount rows
    AnalyzeStats* host = nullptr;
    AnalyzeStats* host_dev = nullptr;  // host's address for the device
};
std::mutex g_an_mu;
std::vector<std::unique_ptr<AnalyzeState>> g_an;

AnalyzeState* analyze_state(int dev) {
    std::lock_guard<std::mutex> lk(g_an_mu);
    if ((int)g_an.size() <= dev) g_an.resize(dev + 1);
    if (!g_an[dev]) g_an[dev].reset(new AnalyzeState);
    return g_an[dev].get();
}

int ensure(AnalyzeState* S) {
    if (!S->rows && hipMalloc((void**)&S->rows, sizeof(AnalyzeRow) * kAnalyzeRowCount) != hipSuccess) {
        S->rows = nullptr;
        return kInternal;
    }
    if (!S->host) {
        void* h = nullptr;
        if (hipHostMalloc(&h, sizeof(AnalyzeStats), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
            return kInternal;
        void* hd = nullptr;
        if (hipHostGetDevicePointer(&hd, h, 0) != hipSuccess) {
            (void)hipHostFree(h);
            return kInternal;
        }
        S->host = (AnalyzeStats*)h;
        S->host_dev = (AnalyzeStats*)hd;
    }
    return kOk;
}

// The statistics are complete in S->host once the stream has drained.
int wait_stats(hipStream_t st) { return hipStreamSynchronize(st) == hipSuccess ? kOk : kInternal; }

}  // namespace
}  // namespace imop_hip

using namespace imop_hip;

extern "C" {

int imop_hip_analyze_data(const void* d_data, uint64_t byte_size, imop_hip_data_info* info, void* stream) {
    if (device_count_cached() == 0) return kInternal;
    if (!info) return kNull;
    std::memset(info, 0, sizeof *info);
    if (!d_data || byte_size == 0) return kOk;  // DataInfo{} (analyzer.cpp:153)
    if (byte_size > 0x7FFFFFFFull) return kInternal;  // the reference's int pixel_count
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return kInternal;
    AnalyzeState* S = analyze_state(dev);
    std::lock_guard<std::mutex> lk(S->mu);
    if (ensure(S) != kOk) return kInternal;
    const hipStream_t st = (hipStream_t)stream;
    const uint64_t n = byte_size, c16 = n / 2u;
    AnalyzeParams p;
    // packed sample counts at depths 10..16 (analyzer.cpp:175-181)
    for (int k = 0; k < 4; ++k) {
        const int bd = 10 + 2 * k;
        uint64_t pc = n * 8u / (uint64_t)bd;
        if (!common_resolution(pc) && common_resolution(pc + 1)) pc++;
        p.sample[k] = (uint32_t)(pc < 10000u ? pc : 10000u);
    }
    const uint64_t n1 = c16 < 100000u ? c16 : 100000u, n2 = c16 < 10000u ? c16 : 10000u;
    // With a known size at n or n / 2 the branch taken reads the whole
    // buffer; otherwise only if detect_bit_depth falls back to 8 bits, which
    // the first launch's statistics tell.
    const bool known = common_resolution(n) || common_resolution(c16);
    if (launch_analyze_stats((const uint8_t*)d_data, n, n1, n2, p, known, S->rows, S->host_dev, st) != hipSuccess ||
        wait_stats(st) != kOk)
        return kInternal;
    const AnalyzeStats& s = *S->host;
    // detect_bit_depth (analyzer.cpp:72-117)
    int bd = 8;
    if (common_resolution(n)) bd = 8;
    else if (n % 2 == 0 && common_resolution(n / 2)) {
        const uint32_t mx = s.max16_n1;
        bd = mx <= 255 ? 8 : mx <= 1023 ? 10 : mx <= 4095 ? 12 : mx <= 16383 ? 14 : 16;
    } else {
        bool decided = false;
        const uint64_t c2 = c16 < 10000u ? c16 : 10000u;
        if (n >= 8 && (double)s.hi_n2 / (double)c2 > 0.85 && (double)s.zero_n2 / (double)c2 < 0.01) {
            for (int d = 10; d <= 16 && !decided; d += 2) {
                const uint64_t pc = n * 8u / (uint64_t)d;
                if (common_resolution(pc) || common_resolution(pc + 1)) {
                    bd = d;
                    decided = true;
                }
            }
        }
        if (!decided && n % 2 == 0 && n >= 4 && (double)s.hi_n2 / (double)c2 > 0.3 && s.max16_n2 > 255) bd = 16;
    }
    // analyze_data (analyzer.cpp:151-214)
    const bool whole = bd <= 8 || common_resolution(c16);
    if (whole && !s.scanned &&
        (launch_analyze_stats((const uint8_t*)d_data, n, n1, n2, p, true, S->rows, S->host_dev, st) != hipSuccess ||
         wait_stats(st) != kOk || !s.scanned))
        return kInternal;
    int pc = 0, mn = 0, mx = 0, packed = 0;
    if (bd > 8) {
        if (common_resolution(c16)) {
            pc = (int)c16;
            mn = (int)s.min16_all;
            mx = (int)s.max16_all;
        } else {
            packed = 1;
            const int k = (bd - 10) / 2;
            uint64_t q = n * 8u / (uint64_t)bd;
            if (!common_resolution(q) && common_resolution(q + 1)) q++;
            pc = (int)q;
            // the reference starts min at safe_max_val and max at 0; a decoded
            // sample never exceeds safe_max_val, so with none decoded min stays there
            mn = (int)std::min<uint32_t>(s.pmin[k], (1u << bd) - 1u);
            mx = (int)s.pmax[k];
        }
    } else {
        pc = (int)n;
        mn = (int)s.min8_all;
        mx = (int)s.max8_all;
    }
    info->detected_bit_depth = bd;
    info->pixel_count = pc;
    info->max_value = mx;
    info->min_value = mn;
    info->is_likely_16bit = bd > 8;
    info->is_packed = packed;
    const auto dims = suggest_dimensions((size_t)pc);
    info->n_dimensions = (int32_t)(dims.size() < 8 ? dims.size() : 8);
    for (int k = 0; k < info->n_dimensions; ++k) {
        info->dimensions[k][0] = dims[k].first;
        info->dimensions[k][1] = dims[k].second;
    }
    if (!dims.empty()) {
        info->suggested_width = dims[0].first;
        info->suggested_height = dims[0].second;
    }
    return kOk;
}

int imop_hip_guess_pattern(const void* d_data, uint64_t byte_size, int width, int height, int bit_depth, int is_packed,
                           int* pattern, void* stream) {
    if (device_count_cached() == 0) return kInternal;
    if (!pattern) return kNull;
    *pattern = 0;  // RGGB (analyzer.cpp:218)
    if (!d_data || width < 4 || height < 4) return kOk;
    if (bit_depth < 1 || bit_depth > 16) return kBitDepth;
    // is_packed as every other entry point takes it: 0, 1, or 2 (GigE Vision
    // 10-bit packing, bit depth 10 only)
    if (is_packed < 0 || is_packed > 2) return kInternal;
    if (is_packed == 2 && bit_depth != 10) return kBitDepth;
    if ((uint64_t)width * (uint64_t)height >= 0x7FFFFFFFull) return kInternal;
    // depths up to 8 are read one byte per pixel, packed or not (analyzer.cpp:228)
    const int packed = bit_depth <= 8 || !is_packed ? 0 : is_packed;
    if (byte_size < bayer_bytes(width, height, bit_depth, packed)) return kInternal;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return kInternal;
    AnalyzeState* S = analyze_state(dev);
    std::lock_guard<std::mutex> lk(S->mu);
    if (ensure(S) != kOk) return kInternal;
    int mode = -1;
    const FrameParams f = make_params(d_data, byte_size, nullptr, width, height, 0, bit_depth, packed, nullptr, 0,
                                      kMg, &mode);
    const long long npx = (long long)width * height;
    const int samples = npx < 20000 ? (int)npx : 20000;
    // analyzer.cpp:242-245, 284-285
    int step = std::max(1, (int)std::sqrt((double)width * height / 5000));
    step = (step / 2) * 2;
    if (step < 2) step = 2;
    int gs = std::max(1, std::min(width, height) / 256);
    gs |= 1;
    const hipStream_t st = (hipStream_t)stream;
    if (launch_guess_pattern_stats(f, samples, step, gs, S->rows, S->host_dev, st) != hipSuccess ||
        wait_stats(st) != kOk)
        return kInternal;
    const AnalyzeStats& s = *S->host;
    // analyzer.cpp:300-318
    const int v1 = (int)s.vote_rgbg, v2 = (int)s.vote_grgb;
    const int margin = v1 > v2 ? v1 - v2 : v2 - v1;
    if (margin <= 2 && v1 + v2 < 10) {
        *pattern = 0;
    } else if (v1 >= v2) {
        const double ee = s.par_cnt[0] ? (double)s.par_sum[0] / s.par_cnt[0] : 0;
        const double oo = s.par_cnt[1] ? (double)s.par_sum[1] / s.par_cnt[1] : 0;
        const double r = ee > 0 ? oo / ee : 0;
        *pattern = r > 1.15 ? 1 : 0;
    } else {
        const double eo = s.par_cnt[2] ? (double)s.par_sum[2] / s.par_cnt[2] : 0;
        const double oe = s.par_cnt[3] ? (double)s.par_sum[3] / s.par_cnt[3] : 0;
        const double r = eo > 0 ? oe / eo : 0;
        *pattern = r > 1.15 ? 3 : 2;
    }
    return kOk;
}

}  // extern "C"
