// imop_batch.cpp -- batched demosaic over frames of mixed size, layout and
// bit depth (imop_hip_demosaic_batch) and its graph-captured plan form.
//
// Frames the register-resident kernel takes (HQLI / MG, even width, u8 / u16
// / 12-bit / 10-bit packed, integral black level, aligned buffers) are
// grouped by (pattern, algorithm); each group becomes one device work list
// and one persistent k_batch launch.  Every other frame gets its single-frame
// launches on the same stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "batch.hpp"
#include "host_common.hpp"
#include "imop_hip.h"
#include "plain.hpp"

namespace imop_hip {

#define IMOP_BATCH_PATTERN_DECL(k)                                                                             \
    hipError_t launch_batch_pattern##k(const BatchArgs& a, int alg, int fmt, bool odd, int rows, bool persist, \
                                       int blocks, hipStream_t st);
IMOP_BATCH_PATTERN_DECL(0)
IMOP_BATCH_PATTERN_DECL(1)
IMOP_BATCH_PATTERN_DECL(2)
IMOP_BATCH_PATTERN_DECL(3)

hipError_t launch_batch(const BatchArgs& a, int pattern, int alg, int fmt, bool odd, int rows, bool persist,
                        int blocks, hipStream_t st) {
    switch (pattern) {
    case 0: return launch_batch_pattern0(a, alg, fmt, odd, rows, persist, blocks, st);
    case 1: return launch_batch_pattern1(a, alg, fmt, odd, rows, persist, blocks, st);
    case 2: return launch_batch_pattern2(a, alg, fmt, odd, rows, persist, blocks, st);
    default: return launch_batch_pattern3(a, alg, fmt, odd, rows, persist, blocks, st);
    }
}

int out_elem_of(const imop_hip_frame_desc& d) {
    return d.out_elem_size == 0 ? (d.bit_depth > 8 ? 2 : 1) : d.out_elem_size;
}

// imop_hip_demosaic's checks, per frame, plus the output element choice.
int check_frame(const imop_hip_frame_desc& d) {
    int e = check_args(d.in, d.out, d.width, d.height, d.pattern, d.algorithm, d.bit_depth, d.is_packed,
                       d.black_level, d.black_level_mode);
    if (e != kOk) return e;
    const int elem = out_elem_of(d);
    if (elem != 1 && elem != 2) return kInternal;
    if (elem == 1 && d.bit_depth > 8) return kInternal;
    const uint64_t px = (uint64_t)d.width * (uint64_t)d.height;
    if (d.in_bytes < bayer_bytes(d.width, d.height, d.bit_depth, d.is_packed) || d.out_bytes < px * 3u * elem)
        return kInternal;
    // u16 samples and u16 RGB are accessed as naturally aligned 16-bit words
    if ((d.bit_depth > 8 && !d.is_packed && ((uintptr_t)d.in & 1)) || (elem == 2 && ((uintptr_t)d.out & 1)))
        return kInternal;
    return kOk;
}

namespace {

// Launch shape of k_batch (imop_hip_set_batch_shape): waves per workgroup
// (4: group items of 4 waves x rows, rows 2 or 4; 1: wave items of rows, 4 or
// 8), workgroups per CU of a persistent grid (0: one workgroup per item),
// order (bit 1: strip-fastest within a frame; bit 2: frame-major placement,
// one-workgroup-per-item grids only).
std::atomic<int> g_rows{4}, g_wpb{4}, g_bpc{0}, g_order{6};

// Rows per wave the form runs: group items take 2 or 4, wave items 4 or 8.
int batch_rows(int rows, int wpb) { return wpb == 1 ? (rows == 8 ? 8 : 4) : (rows == 2 ? 2 : 4); }

int cu_count() {
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
        cache[dev] = n;
    }
    return cache[dev];
}

// One launch: the frames of one (pattern, algorithm, raw layout, width class).
struct Group {
    int pattern = 0, alg = 0, fmt = 0;
    bool odd = false;
    std::vector<BatchFrame> frames;
    uint32_t total = 0;           // items, or slots under frame-major placement (BatchFrame::first likewise)
    std::vector<uint32_t> chunk;  // {frame of entry c * kBatchChunk, first entry of the next frame} by chunk c

    // The device copy: the frame table, then the chunk table at a 16-byte boundary.
    size_t chunk_offset() const { return (frames.size() * sizeof(BatchFrame) + 15u) & ~(size_t)15u; }
    size_t device_bytes() const { return chunk_offset() + chunk.size() * sizeof(uint32_t); }
    std::vector<uint8_t> device_image() const {
        std::vector<uint8_t> b(device_bytes(), 0);
        std::memcpy(b.data(), frames.data(), frames.size() * sizeof(BatchFrame));
        std::memcpy(b.data() + chunk_offset(), chunk.data(), chunk.size() * sizeof(uint32_t));
        return b;
    }
};

struct Single {
    FrameParams f;
    int algo, mode;
};

struct Prepared {
    std::vector<Group> groups;
    std::vector<Single> singles;
    size_t prism_ws = 0;
    int rows = 4, wpb = 4, bpc = 0, order = 6;
};

int prepare(const imop_hip_frame_desc* fr, int n, Prepared& P, int* failed) {
    if (failed) *failed = -1;
    if (n < 0) return kInternal;
    if (n > 0 && !fr) return kNull;
    for (int i = 0; i < n; ++i) {
        const int e = check_frame(fr[i]);
        if (e != kOk) {
            if (failed) *failed = i;
            return e;
        }
    }
    P.wpb = g_wpb;
    P.rows = batch_rows(g_rows, P.wpb);
    P.bpc = g_bpc;
    // frame-major placement is a form of the one-workgroup-per-item grid
    P.order = g_order & (P.bpc > 0 ? 2 : 6);
    const bool frame_major = (P.order & 4) != 0;
    std::map<int, size_t> index;  // group of each launch key
    for (int i = 0; i < n; ++i) {
        const imop_hip_frame_desc& d = fr[i];
        int mode = -1;
        FrameParams f = make_params(d.in, d.in_bytes, d.out, d.width, d.height, d.pattern, d.bit_depth, d.is_packed,
                                    d.black_level, d.black_level_mode, d.algorithm, &mode);
        f.out16 = out_elem_of(d) == 2;
        // the single-frame kernel's items: aligned strips for u16 RGB by
        // group rows of rows x 4 (group form) or bands of rows (wave form)
        const int nsx = strip_count(d.width, f.out16 != 0);
        const int nby = (d.height + P.rows * P.wpb - 1) / (P.rows * P.wpb);
        // k_batch runs MG 8-bit frames on the integer path; a packed 8-bit
        // frame without black level takes the exact path, so it runs alone
        const bool mode_ok = d.algorithm == kHqli || (d.algorithm == kMg && (f.fmt == FMT_U8) == (mode == MODE_MG_INT8));
        // k_batch decodes u8 / u16 / 12-bit / 10-bit reference-packed runs only
        const bool batched = (d.algorithm == kHqli || d.algorithm == kMg) && mode_ok && fast_path_ok(f) &&
                             f.fmt != FMT_G10 &&
                             nsx <= 65535 && nby <= 65535;
        if (!batched) {
            P.singles.push_back(Single{f, d.algorithm, mode});
            if (d.algorithm == kPrism || d.algorithm == kRcd) P.prism_ws = std::max(P.prism_ws, prism_workspace_bytes(d.width, d.height));
            continue;
        }
        const int alg = d.algorithm == kMg ? 1 : 0, odd = (d.width & 3) == 2 ? 1 : 0;
        const int key = ((d.pattern * 2 + alg) * 8 + f.fmt) * 2 + odd;
        auto it = index.find(key);
        if (it == index.end()) {
            it = index.emplace(key, P.groups.size()).first;
            P.groups.emplace_back();
            P.groups.back().pattern = d.pattern;
            P.groups.back().alg = alg;
            P.groups.back().fmt = f.fmt;
            P.groups.back().odd = odd != 0;
        }
        Group& g = P.groups[it->second];
        const uint64_t items = (uint64_t)nsx * (uint64_t)nby;
        // 31-bit item numbers; at most 65535 frames per launch
        if ((uint64_t)g.total + items >= (1ull << 31) || g.frames.size() >= 65535u) return kInternal;
        BatchFrame b{};
        b.in = f.in;
        b.out = f.out;
        b.in_bytes = (uint32_t)f.in_bytes;
        b.w = f.w;
        b.h = f.h;
        b.bd = f.bd;
        b.maxv = f.maxv;
        b.clamp_hi = f.clamp_hi;
        b.fmt = f.fmt;
        b.flags = (f.out16 ? 1 : 0) | (f.bl_on ? 2 : 0);
        for (int k = 0; k < 4; ++k) b.bl_i[k] = f.bl_on ? (int)f.bl_q[k] : 0;
        b.nsx = (uint16_t)nsx;
        b.nby = (uint16_t)nby;
        b.first = g.total;
        b.place_mul = place_reciprocal(items, (P.order & 2) ? (uint64_t)nsx : (uint64_t)nby);
        g.frames.push_back(b);
        // frame-major: the frame takes ceil(items / 8) slots of 8 workgroups
        g.total += frame_major ? (uint32_t)((items + 7) / 8) : (uint32_t)items;
    }
    for (const Group& g : P.groups) {
        // grid of one workgroup per item (slot) fits the launch limits
        const uint64_t blocks = frame_major ? 8ull * g.total : (uint64_t)g.total;
        if (P.bpc == 0 && blocks * (P.wpb == 1 ? 64u : 256u) >= (1ull << 32)) return kInternal;
    }
    for (Group& g : P.groups) {
        const uint32_t nchunks = (g.total + kBatchChunk - 1) / kBatchChunk;
        g.chunk.resize(2 * (size_t)nchunks);
        size_t fi = 0;
        for (uint32_t c = 0; c < nchunks; ++c) {
            const uint32_t item = c * (uint32_t)kBatchChunk;  // an item, or a slot
            while (fi + 1 < g.frames.size() && g.frames[fi + 1].first <= item) ++fi;
            g.chunk[2 * c] = (uint32_t)fi;
            g.chunk[2 * c + 1] = fi + 1 < g.frames.size() ? g.frames[fi + 1].first : g.total;
        }
    }
    return kOk;
}

// Enqueue every launch of a prepared batch; tables[i] = device copy of
// groups[i] (Group::device_image).
int enqueue(const Prepared& P, const std::vector<void*>& tables, void* prism_ws, hipStream_t st) {
    const int cus = cu_count();
    for (size_t i = 0; i < P.groups.size(); ++i) {
        const Group& g = P.groups[i];
        if (g.total == 0) continue;
        BatchArgs a;
        a.frames = (const BatchFrame*)tables[i];
        a.chunk = (const uint2*)((const uint8_t*)tables[i] + g.chunk_offset());
        a.nframes = (int)g.frames.size();
        a.total = g.total;
        a.wpb = P.wpb;
        a.order = P.order;
        // one workgroup per item, or 8 per slot (frame-major)
        int64_t blocks = (P.order & 4) ? 8 * (int64_t)g.total : (int64_t)g.total;
        if (P.bpc > 0 && cus > 0) blocks = std::min<int64_t>(blocks, (int64_t)P.bpc * cus);
        if (launch_batch(a, g.pattern, g.alg, g.fmt, g.odd, P.rows, P.bpc > 0, (int)blocks, st) != hipSuccess)
            return kInternal;
    }
    for (const Single& s : P.singles) {
        const int e = launch_frame(s.f, s.algo, s.mode, prism_ws, st);
        if (e != kOk) return e;
    }
    return kOk;
}

}  // namespace

}  // namespace imop_hip

using namespace imop_hip;

struct imop_hip_batch_plan {
    int dev = 0;
    std::vector<void*> tables;
    void* prism_ws = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void free_plan(imop_hip_batch_plan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (void* t : p->tables)
        if (t) (void)hipFree(t);
    if (p->prism_ws) (void)hipFree(p->prism_ws);
    delete p;
}

}  // namespace

extern "C" {

int imop_hip_demosaic_batch(const imop_hip_frame_desc* frames, int n, void* stream, int* failed_frame) {
    if (failed_frame) *failed_frame = -1;
    if (device_count_cached() == 0) return kInternal;
    Prepared P;
    int e = prepare(frames, n, P, failed_frame);
    if (e != kOk) return e;
    hipStream_t st = (hipStream_t)stream;
    // The work lists are copied from this call's host memory, which a graph
    // replay would read after the call returned: graphs take the plan form.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return kInternal;
    // every launch's work list in one stream-ordered allocation (one copy,
    // each list at a 256-byte boundary), released after the launches
    std::vector<size_t> offset(P.groups.size(), 0);
    size_t bytes = 0;
    for (size_t i = 0; i < P.groups.size(); ++i) {
        offset[i] = bytes;
        bytes += (P.groups[i].device_bytes() + 255u) & ~(size_t)255u;
    }
    std::vector<void*> tables(P.groups.size(), nullptr);
    if (bytes == 0) return enqueue(P, tables, nullptr, st);
    // kept alive until the call returns: the copy reads it in stream order
    std::vector<uint8_t> image(bytes, 0);
    for (size_t i = 0; i < P.groups.size(); ++i) {
        const std::vector<uint8_t> g = P.groups[i].device_image();
        std::memcpy(image.data() + offset[i], g.data(), g.size());
    }
    void* base = nullptr;
    if (hipMallocAsync(&base, bytes, st) != hipSuccess) return kInternal;
    if (hipMemcpyAsync(base, image.data(), bytes, hipMemcpyHostToDevice, st) != hipSuccess) e = kInternal;
    for (size_t i = 0; i < P.groups.size(); ++i) tables[i] = (uint8_t*)base + offset[i];
    if (e == kOk) e = enqueue(P, tables, nullptr, st);
    (void)hipFreeAsync(base, st);
    return e;
}

int imop_hip_batch_plan_create(const imop_hip_frame_desc* frames, int n, imop_hip_batch_plan** plan,
                               int* failed_frame) {
    if (failed_frame) *failed_frame = -1;
    if (!plan) return kNull;
    *plan = nullptr;
    if (device_count_cached() == 0) return kInternal;
    Prepared P;
    int e = prepare(frames, n, P, failed_frame);
    if (e != kOk) return e;
    imop_hip_batch_plan* p = new imop_hip_batch_plan();
    if (hipGetDevice(&p->dev) != hipSuccess) {
        free_plan(p);
        return kInternal;
    }
    p->tables.assign(P.groups.size(), nullptr);
    for (size_t i = 0; i < P.groups.size(); ++i) {
        const std::vector<uint8_t> image = P.groups[i].device_image();
        if (hipMalloc(&p->tables[i], image.size()) != hipSuccess ||
            hipMemcpy(p->tables[i], image.data(), image.size(), hipMemcpyHostToDevice) != hipSuccess) {
            free_plan(p);
            return kInternal;
        }
    }
    if (P.prism_ws && hipMalloc(&p->prism_ws, P.prism_ws) != hipSuccess) {
        free_plan(p);
        return kInternal;
    }
    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
        free_plan(p);
        return kInternal;
    }
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipStreamDestroy(cs);
        free_plan(p);
        return kInternal;
    }
    e = enqueue(P, p->tables, p->prism_ws, cs);
    const hipError_t ce = hipStreamEndCapture(cs, &p->graph);
    (void)hipStreamDestroy(cs);
    if (e != kOk || ce != hipSuccess ||
        hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess) {
        free_plan(p);
        return kInternal;
    }
    *plan = p;
    return kOk;
}

int imop_hip_batch_plan_launch(imop_hip_batch_plan* plan, void* stream) {
    if (!plan) return kNull;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev != plan->dev) return kInternal;
    return hipGraphLaunch(plan->exec, (hipStream_t)stream) == hipSuccess ? kOk : kInternal;
}

int imop_hip_batch_plan_destroy(imop_hip_batch_plan* plan) {
    if (!plan) return kNull;
    free_plan(plan);
    return kOk;
}

int imop_hip_set_batch_shape(int rows_per_wave, int waves_per_block, int blocks_per_cu, int order) {
    if (rows_per_wave == 2 || rows_per_wave == 4 || rows_per_wave == 8) g_rows = rows_per_wave;
    if (waves_per_block == 1 || waves_per_block == 4) g_wpb = waves_per_block;
    if (blocks_per_cu >= 0 && blocks_per_cu <= 64) g_bpc = blocks_per_cu;
    if (order >= 0 && order <= 7) g_order = order & 6;
    return kOk;
}

int imop_hip_get_batch_shape(int* rows_per_wave, int* waves_per_block, int* blocks_per_cu, int* order) {
    if (rows_per_wave) *rows_per_wave = g_rows;
    if (waves_per_block) *waves_per_block = g_wpb;
    if (blocks_per_cu) *blocks_per_cu = g_bpc;
    if (order) *order = g_order;
    return kOk;
}

}  // extern "C"
