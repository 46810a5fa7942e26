// imop_stream.cpp -- demosaic of frames in pageable host memory through
// pinned, triple-buffered staging (imop_hip_demosaic_batch_host).
//
// Frames are grouped in order into chunks of about kChunkBytes (input +
// output).  Chunk c uses staging slot s = c % kSlots: a pinned host input and
// output area, a device input and output area, and stream s.  On stream s:
//   H2D(c) -> wait for kernel(c-1) -> imop_hip_demosaic_batch(c) -> D2H(c).
// Kernels of consecutive chunks are serialised through that wait (one chunk's
// kernels already fill the device), while the copies of neighbouring chunks
// overlap them.  On the host, the caller's thread scatters chunk c - kSlots's outputs
// and gathers chunk c's inputs (threaded memcpy) while the GPU works on the
// chunks in between: with three slots, two chunks' copies stay queued while
// the host handles a large frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "host_common.hpp"
#include "imop_hip.h"

namespace imop_hip {
namespace {

constexpr size_t kChunkBytes = (size_t)96 << 20;  // input + output staged per chunk (a larger frame: its own)
constexpr int kSlots = 3;
constexpr size_t kParallelCopy = (size_t)4 << 20;  // below this, one thread copies

size_t align256(size_t v) { return (v + 255u) & ~(size_t)255u; }

struct Seg {
    uint8_t* dst;
    const uint8_t* src;
    size_t n;
};

// Threads of the host-side gathers and scatters: IMOP_STREAM_THREADS when
// set (1..64), else the hardware threads, at most 16.
int copy_threads() {
    static const int n = [] {
        const char* e = std::getenv("IMOP_STREAM_THREADS");
        if (e && std::atoi(e) > 0) return std::min(64, std::atoi(e));
        const unsigned hw = std::thread::hardware_concurrency();
        return (int)std::max(1u, std::min(16u, hw ? hw : 1u));
    }();
    return n;
}

// Copy every segment, split into byte ranges over copy_threads() threads.
void par_copy(const std::vector<Seg>& segs) {
    size_t total = 0;
    for (const Seg& s : segs) total += s.n;
    if (total == 0) return;
    const int nt = total < kParallelCopy ? 1 : copy_threads();
    auto run = [&](size_t lo, size_t hi) {  // bytes [lo, hi) of the concatenated segments
        size_t base = 0;
        for (const Seg& s : segs) {
            const size_t a = std::max(lo, base), b = std::min(hi, base + s.n);
            if (a < b) std::memcpy(s.dst + (a - base), s.src + (a - base), b - a);
            base += s.n;
            if (base >= hi) break;
        }
    };
    if (nt == 1) {
        run(0, total);
        return;
    }
    const size_t per = ((total + nt - 1) / nt + 4095u) & ~(size_t)4095u;
    std::vector<std::thread> th;
    size_t done_to = per;  // ranges [per, done_to) have a thread; the caller copies the rest
    try {
        th.reserve(nt - 1);
        for (int t = 1; t < nt; ++t) {
            const size_t lo = std::min(total, per * t), hi = std::min(total, per * (t + 1));
            if (lo < hi) th.emplace_back(run, lo, hi);
            done_to = std::max(done_to, hi);
        }
    } catch (...) {  // no more threads: nothing may escape through the C ABI
    }
    run(0, std::min(total, per));
    if (done_to < total) run(done_to, total);
    for (std::thread& t : th) t.join();
}

// Per-host-thread staging: kSlots slots of pinned host and device memory,
// one stream each (default priority) and their events.  Grows, never shrinks.
struct StreamStaging {
    int device = -1;
    hipStream_t st[kSlots] = {};
    hipEvent_t kernel_done[kSlots] = {}, done[kSlots] = {};
    uint8_t* h[kSlots] = {};
    uint8_t* d[kSlots] = {};
    size_t cap = 0;

    ~StreamStaging() { release(); }
    void release() {
        for (int s = 0; s < kSlots; ++s) {
            if (st[s]) (void)hipStreamSynchronize(st[s]);
            if (h[s]) (void)hipHostFree(h[s]);
            if (d[s]) (void)hipFree(d[s]);
            if (kernel_done[s]) (void)hipEventDestroy(kernel_done[s]);
            if (done[s]) (void)hipEventDestroy(done[s]);
            if (st[s]) (void)hipStreamDestroy(st[s]);
            h[s] = d[s] = nullptr;
            kernel_done[s] = done[s] = nullptr;
            st[s] = nullptr;
        }
        cap = 0;
        device = -1;
    }
    bool ensure(size_t bytes) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return false;
        if (dev != device) {
            release();
            for (int s = 0; s < kSlots; ++s)
                if (hipStreamCreateWithFlags(&st[s], hipStreamNonBlocking) != hipSuccess ||
                    hipEventCreateWithFlags(&kernel_done[s], hipEventDisableTiming) != hipSuccess ||
                    hipEventCreateWithFlags(&done[s], hipEventDisableTiming) != hipSuccess) {
                    release();
                    return false;
                }
            device = dev;
        }
        if (cap >= bytes) return true;
        for (int s = 0; s < kSlots; ++s) {
            if (h[s]) (void)hipHostFree(h[s]);
            if (d[s]) (void)hipFree(d[s]);
            h[s] = d[s] = nullptr;
        }
        cap = 0;
        for (int s = 0; s < kSlots; ++s)
            if (hipHostMalloc((void**)&h[s], bytes, hipHostMallocDefault) != hipSuccess ||
                hipMalloc((void**)&d[s], bytes) != hipSuccess)
                return false;
        cap = bytes;
        return true;
    }
};
thread_local StreamStaging t_stream;

struct Placed {
    size_t in_off, in_n, out_off, out_n;  // within the slot; bytes staged / returned
};

struct Chunk {
    int first = 0, count = 0;
    size_t in_total = 0, bytes = 0;  // inputs occupy [0, in_total); outputs follow
    std::vector<Placed> at;
};

}  // namespace
}  // namespace imop_hip

using namespace imop_hip;

extern "C" int imop_hip_demosaic_batch_host(const imop_hip_frame_desc* frames, int n, int* failed_frame) {
    if (failed_frame) *failed_frame = -1;
    if (device_count_cached() == 0) return kInternal;
    if (n < 0) return kInternal;
    if (n == 0) return kOk;
    if (!frames) return kNull;
    // Every frame is checked before anything is copied.  Host buffers need no
    // alignment: they are staged at 256-byte offsets.
    std::vector<size_t> in_n(n), out_n(n);
    for (int i = 0; i < n; ++i) {
        imop_hip_frame_desc c = frames[i];
        if (c.in) c.in = (const void*)(uintptr_t)256;
        if (c.out) c.out = (void*)(uintptr_t)256;
        const int e = check_frame(c);
        if (e != kOk) {
            if (failed_frame) *failed_frame = i;
            return e;
        }
        in_n[i] = bayer_bytes(c.width, c.height, c.bit_depth, c.is_packed);
        out_n[i] = (size_t)c.width * (size_t)c.height * 3u * (size_t)out_elem_of(c);
    }
    // chunks in frame order
    std::vector<Chunk> chunks;
    size_t need = 0;
    for (int i = 0; i < n;) {
        Chunk ch;
        ch.first = i;
        size_t in_sum = 0, out_sum = 0;
        while (i < n) {
            const size_t a = align256(in_n[i]), b = align256(out_n[i]);
            if (ch.count > 0 && in_sum + out_sum + a + b > kChunkBytes) break;
            in_sum += a;
            out_sum += b;
            ++ch.count;
            ++i;
        }
        ch.in_total = in_sum;
        ch.bytes = in_sum + out_sum;
        size_t io = 0, oo = in_sum;
        for (int k = 0; k < ch.count; ++k) {
            const int f = ch.first + k;
            ch.at.push_back({io, in_n[f], oo, out_n[f]});
            io += align256(in_n[f]);
            oo += align256(out_n[f]);
        }
        need = std::max(need, ch.bytes);
        chunks.push_back(std::move(ch));
    }
    StreamStaging& S = t_stream;
    if (!S.ensure(need)) return kInternal;

    auto gather = [&](const Chunk& ch, int s) {
        std::vector<Seg> segs;
        for (int k = 0; k < ch.count; ++k)
            segs.push_back({S.h[s] + ch.at[k].in_off, (const uint8_t*)frames[ch.first + k].in, ch.at[k].in_n});
        par_copy(segs);
    };
    auto scatter = [&](const Chunk& ch, int s) {
        std::vector<Seg> segs;
        for (int k = 0; k < ch.count; ++k)
            segs.push_back({(uint8_t*)frames[ch.first + k].out, S.h[s] + ch.at[k].out_off, ch.at[k].out_n});
        par_copy(segs);
    };
    auto fail = [&](int e) {
        for (int s = 0; s < kSlots; ++s) (void)hipStreamSynchronize(S.st[s]);
        return e;
    };

    std::vector<imop_hip_frame_desc> descs;
    const int nc = (int)chunks.size();
    for (int c = 0; c < nc; ++c) {
        const int s = c % kSlots;
        const Chunk& ch = chunks[c];
        if (c >= kSlots) {  // slot s last held chunk c - kSlots
            if (hipEventSynchronize(S.done[s]) != hipSuccess) return fail(kInternal);
            scatter(chunks[c - kSlots], s);
        }
        gather(ch, s);
        if (hipMemcpyAsync(S.d[s], S.h[s], ch.in_total, hipMemcpyHostToDevice, S.st[s]) != hipSuccess)
            return fail(kInternal);
        if (c >= 1 && hipStreamWaitEvent(S.st[s], S.kernel_done[(c - 1) % kSlots], 0) != hipSuccess)
            return fail(kInternal);
        descs.assign(frames + ch.first, frames + ch.first + ch.count);
        for (int k = 0; k < ch.count; ++k) {
            descs[k].in = S.d[s] + ch.at[k].in_off;
            descs[k].in_bytes = ch.at[k].in_n;
            descs[k].out = S.d[s] + ch.at[k].out_off;
            descs[k].out_bytes = ch.at[k].out_n;
        }
        int bad = -1;
        const int e = imop_hip_demosaic_batch(descs.data(), ch.count, S.st[s], &bad);
        if (e != kOk) {
            if (failed_frame && bad >= 0) *failed_frame = ch.first + bad;
            return fail(e);
        }
        if (hipEventRecord(S.kernel_done[s], S.st[s]) != hipSuccess ||
            hipMemcpyAsync(S.h[s] + ch.in_total, S.d[s] + ch.in_total, ch.bytes - ch.in_total,
                           hipMemcpyDeviceToHost, S.st[s]) != hipSuccess ||
            hipEventRecord(S.done[s], S.st[s]) != hipSuccess)
            return fail(kInternal);
    }
    for (int c = std::max(0, nc - kSlots); c < nc; ++c) {
        const int s = c % kSlots;
        if (hipEventSynchronize(S.done[s]) != hipSuccess) return fail(kInternal);
        scatter(chunks[c], s);
    }
    return kOk;
}
