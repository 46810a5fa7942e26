// analyze_kernels.hip -- the statistics behind imop::analyze_data and
// imop::guess_pattern (src/analyzer.cpp:13-321), gathered on the GPU from a
// raw buffer already in device memory.  Every statistic is a maximum, a
// minimum, a count or a sum of integers (exact in any order), so the host can
// take the reference's decisions on them and reach the same result.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "analyze.hpp"
#include "device_common.hpp"

namespace imop_hip {
namespace {

constexpr int kThreads = 256;
enum : uint32_t { OP_SUM = 0, OP_MIN = 1, OP_MAX = 2 };

__device__ __forceinline__ unsigned long long comb(uint32_t op, unsigned long long a, unsigned long long b) {
    return op == OP_SUM ? a + b : op == OP_MIN ? (a < b ? a : b) : (a > b ? a : b);
}

// Reduce v over the workgroup (kThreads lanes) with the k-th value's
// operation in bits 2k..2k+1 of ops; threads 0..N-1 write the results to
// row[0..N-1].  Every thread of the workgroup must call it.
template <int N>
__device__ __forceinline__ void block_reduce(unsigned long long (&v)[N], uint32_t ops, unsigned long long* row) {
    __shared__ unsigned long long sm[kThreads / 64][N];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const uint32_t op = (ops >> (2 * k)) & 3u;
        for (int o = 32; o > 0; o >>= 1) v[k] = comb(op, v[k], __shfl_xor(v[k], o));
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) sm[wv][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < N) {
        const int k = threadIdx.x;
        const uint32_t op = (ops >> (2 * k)) & 3u;
        unsigned long long r = sm[0][k];
        for (int w = 1; w < kThreads / 64; ++w) r = comb(op, r, sm[w][k]);
        row[k] = r;
    }
    __syncthreads();
}

constexpr unsigned long long kNone = ~0ull;  // neutral for OP_MIN

// The first N values of rows[0, nrows) reduced over the workgroup into out
// (LDS), visible to every thread on return.
template <int N>
__device__ __forceinline__ void reduce_rows(const AnalyzeRow* rows, int nrows, uint32_t ops,
                                            unsigned long long* out) {
    unsigned long long v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = ((ops >> (2 * k)) & 3u) == OP_MIN ? kNone : 0ull;
    for (int q = threadIdx.x; q < nrows; q += blockDim.x)
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = comb((ops >> (2 * k)) & 3u, v[k], rows[q].v[k]);
    block_reduce(v, ops, out);
}



// Row of k_analyze: 0 max of the first n1 u16; 1 high-byte count, 2 zero
// count and 3 max of the first n2; 4..7 / 8..11 min / max of the packed
// samples at depths 10, 12, 14, 16; 12 / 13 min / max of every u16, 14 / 15
// min / max of every byte.
constexpr uint32_t kAnalyzeOps = OP_MAX | OP_SUM << 2 | OP_SUM << 4 | OP_MAX << 6 | OP_MIN << 8 | OP_MIN << 10 |
                                 OP_MIN << 12 | OP_MIN << 14 | OP_MAX << 16 | OP_MAX << 18 | OP_MAX << 20 |
                                 OP_MAX << 22 | OP_MIN << 24 | OP_MAX << 26 | OP_MIN << 28 | OP_MAX << 30;

// Running minima / maxima of the whole-buffer scan as packed 16-bit pairs:
// the word's two u16 values, and its four bytes as two pairs (even and odd
// bytes widened to 16 bits), one packed min / max instruction each.
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
struct ScanAcc {
    u16x2 mn16, mx16, mn8, mx8;
};

__device__ __forceinline__ void scan_word(uint32_t w, ScanAcc& a) {
    const u16x2 x = __builtin_bit_cast(u16x2, w);
    a.mn16 = __builtin_elementwise_min(a.mn16, x);
    a.mx16 = __builtin_elementwise_max(a.mx16, x);
    const u16x2 e = __builtin_bit_cast(u16x2, w & 0x00FF00FFu), o = __builtin_bit_cast(u16x2, (w >> 8) & 0x00FF00FFu);
    a.mn8 = __builtin_elementwise_min(a.mn8, __builtin_elementwise_min(e, o));
    a.mx8 = __builtin_elementwise_max(a.mx8, __builtin_elementwise_max(e, o));
}

// Steps of kScanU 16-byte chunks per thread, all loads of a step issued first.
constexpr int kScanU = 4;


// All statistics analyze_data reads, one row per workgroup (reduced by
// k_analyze_final):
//   - the prefixes of the bytes d[0, n) read as little-endian u16 (the first
//     n1 and n2 values) and the LSB-first 2-byte-window decode of the first
//     sample[k] pixels at depths 10, 12, 14, 16 (at most 100000 values);
//   - with scan, the whole buffer's range as u16 and as bytes (the unpacked
//     16- and 8-bit branches, analyzer.cpp:163-170, 195-202): 16-byte chunks,
//     kScanU in flight per thread, when d is 16-byte aligned, then the tail
//     value by value.
__global__ __launch_bounds__(kThreads) void k_analyze(const uint8_t* __restrict__ d, uint64_t n, uint64_t n1,
                                                      uint64_t n2, AnalyzeParams p, int scan, AnalyzeRow* rows) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v[16] = {0, 0, 0, 0, kNone, kNone, kNone, kNone, 0, 0, 0, 0, kNone, 0, kNone, 0};
    auto u16 = [&](uint64_t i) { return (uint32_t)d[2 * i] | ((uint32_t)d[2 * i + 1] << 8); };
    for (uint64_t i = t0; i < n1; i += stride) v[0] = max(v[0], (unsigned long long)u16(i));
    for (uint64_t i = t0; i < n2; i += stride) {
        const uint32_t x = u16(i);
        v[1] += (x >> 8) != 0u;
        v[2] += x == 0u;
        v[3] = max(v[3], (unsigned long long)x);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int bd = 10 + 2 * k;
        for (uint64_t i = t0; i < (uint64_t)p.sample[k]; i += stride) {
            const uint64_t bo = (i * (uint64_t)bd) / 8u, bs = (i * (uint64_t)bd) % 8u;
            if (bo + 1u < n) {
                const uint32_t x = (((uint32_t)d[bo] | ((uint32_t)d[bo + 1] << 8)) >> bs) & ((1u << bd) - 1u);
                v[4 + k] = min(v[4 + k], (unsigned long long)x);
                v[8 + k] = max(v[8 + k], (unsigned long long)x);
            }
        }
    }
    if (scan) {
        const uint64_t count16 = n / 2u;
        const uint64_t nv = ((uintptr_t)d & 15u) == 0 ? n / 16u : 0;
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4* dv = (const u32x4*)d;
        ScanAcc acc = {{0xFFFF, 0xFFFF}, {0, 0}, {0xFFFF, 0xFFFF}, {0, 0}};
        uint64_t c = t0;
        for (; c + (kScanU - 1) * stride < nv; c += kScanU * stride) {
            u32x4 q[kScanU];
#pragma unroll
            for (int u = 0; u < kScanU; ++u) q[u] = __builtin_nontemporal_load(dv + c + u * stride);
#pragma unroll
            for (int u = 0; u < kScanU; ++u) {
                scan_word(q[u].x, acc);
                scan_word(q[u].y, acc);
                scan_word(q[u].z, acc);
                scan_word(q[u].w, acc);
            }
        }
        for (; c < nv; c += stride) {
            const u32x4 a = __builtin_nontemporal_load(dv + c);
            scan_word(a.x, acc);
            scan_word(a.y, acc);
            scan_word(a.z, acc);
            scan_word(a.w, acc);
        }
        if (nv > 0) {  // 0xFFFF / 0 where a thread took no chunk: no value is above / below them
            v[12] = min(v[12], (unsigned long long)min(acc.mn16.x, acc.mn16.y));
            v[13] = max(v[13], (unsigned long long)max(acc.mx16.x, acc.mx16.y));
            v[14] = min(v[14], (unsigned long long)min(acc.mn8.x, acc.mn8.y));
            v[15] = max(v[15], (unsigned long long)max(acc.mx8.x, acc.mx8.y));
        }
        for (uint64_t i = 8u * nv + t0; i < count16; i += stride) {
            const uint32_t x = u16(i);
            v[12] = min(v[12], (unsigned long long)x);
            v[13] = max(v[13], (unsigned long long)x);
            v[14] = min(v[14], (unsigned long long)min(x & 0xFFu, x >> 8));
            v[15] = max(v[15], (unsigned long long)max(x & 0xFFu, x >> 8));
        }
        if (t0 == 0 && (n & 1u)) {  // the odd last byte
            v[14] = min(v[14], (unsigned long long)d[n - 1]);
            v[15] = max(v[15], (unsigned long long)d[n - 1]);
        }
    }
    block_reduce(v, kAnalyzeOps, rows[blockIdx.x].v);
}

// One workgroup: k_analyze's nrows rows into *out (pinned host memory mapped
// for the device).
__global__ __launch_bounds__(kThreads) void k_analyze_final(const AnalyzeRow* rows, int nrows, int scan,
                                                            AnalyzeStats* out) {
    __shared__ unsigned long long tot[16];
    reduce_rows<16>(rows, nrows, kAnalyzeOps, tot);
    if (threadIdx.x == 0) {
        AnalyzeStats z = {};
        z.max16_n1 = (uint32_t)tot[0];
        z.hi_n2 = (uint32_t)tot[1];
        z.zero_n2 = (uint32_t)tot[2];
        z.max16_n2 = (uint32_t)tot[3];
        for (int k = 0; k < 4; ++k) {
            z.pmin[k] = (uint32_t)min(tot[4 + k], 0xFFFFFFFFull);
            z.pmax[k] = (uint32_t)tot[8 + k];
        }
        z.min16_all = (uint32_t)min(tot[12], 0xFFFFFFFFull);
        z.max16_all = (uint32_t)tot[13];
        z.min8_all = (uint32_t)min(tot[14], 0xFFFFFFFFull);
        z.max8_all = (uint32_t)tot[15];
        z.scanned = scan ? 1u : 0u;
        *out = z;
    }
}

// guess_pattern's sample of up to 20000 pixels at (i * 10007) % (w * h), and
// the pixel (0, 0) (analyzer.cpp:230-238): row slots 0 / 1 min / max.
__global__ __launch_bounds__(kThreads) void k_pattern_sample(FrameParams f, int samples, AnalyzeRow* rows) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int npx = f.w * f.h;
    unsigned long long v[2] = {kNone, 0};
    if (i < samples) {
        const int idx = (i * 10007) % npx;
        v[0] = v[1] = (unsigned long long)sample_at(f, idx % f.w, idx / f.w);  // samples are never negative
    }
    if (i == 0) {
        const unsigned long long x = (unsigned long long)sample_at(f, 0, 0);
        v[0] = min(v[0], x);
        v[1] = max(v[1], x);
    }
    block_reduce(v, OP_MIN | OP_MAX << 2, rows[blockIdx.x].v);
}

// Workgroups [0, nvote): the 2x2 block votes (analyzer.cpp:254-282) in
// double, as the reference, into row slots 8 and 9; the others: sums and
// counts of the samples at the four CFA parities over every gs-th row and
// column (analyzer.cpp:284-298), slots 0..3 and 4..7.  All sums (reduced
// by k_pattern_final).  The vote workgroups first reduce k_pattern_sample's
// rows to the sample range.
__global__ __launch_bounds__(kThreads) void k_pattern_blocks(FrameParams f, int step, int nbx, int nblocks, int gs,
                                                             int ngx, int npos, int nvote, const AnalyzeRow* srows,
                                                             int nsrows, AnalyzeRow* rows) {
    unsigned long long v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if ((int)blockIdx.x < nvote) {
        __shared__ unsigned long long mm[2];
        reduce_rows<2>(srows, nsrows, OP_MIN | OP_MAX << 2, mm);
        const int range = max(1, (int)mm[1] - (int)mm[0]);
        double conf = range * 0.02;
        if (conf < 1.0) conf = 1.0;
        for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nblocks; b += nvote * blockDim.x) {
            const int x = (b % nbx) * step, y = (b / nbx) * step;
            const double tl = sample_at(f, x, y), tr = sample_at(f, x + 1, y);
            const double bl = sample_at(f, x, y + 1), br = sample_at(f, x + 1, y + 1);
            const double d1 = fabs(tl - br), d2 = fabs(tr - bl);
            bool ok = !(fabs(d1 - d2) < conf);
            const double hi4 = fmax(fmax(tl, tr), fmax(bl, br)), lo4 = fmin(fmin(tl, tr), fmin(bl, br));
            ok = ok && !(hi4 - lo4 > range * 0.15);
            const double hd = fabs(tl - tr) + fabs(bl - br), vd = fabs(tl - bl) + fabs(tr - br);
            const double mn = hd < vd ? hd : vd, mxd = hd < vd ? vd : hd;
            ok = ok && !(mxd / (mn > 1.0 ? mn : 1.0) > 4.0);
            if (ok) v[d2 < d1 ? 8 : 9] += 1u;
        }
    } else {
        const int nb = gridDim.x - nvote;
        for (int q = (blockIdx.x - nvote) * blockDim.x + threadIdx.x; q < npos; q += nb * blockDim.x) {
            const int x = (q % ngx) * gs, y = (q / ngx) * gs;
            const int k = ((y & 1) == 0 && (x & 1) == 0) ? 0 : ((y & 1) == 1 && (x & 1) == 1) ? 1 : ((y & 1) == 0) ? 2 : 3;
            const unsigned long long x2 = (unsigned long long)sample_at(f, x, y);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                v[j] += j == k ? x2 : 0ull;
                v[4 + j] += j == k ? 1ull : 0ull;
            }
        }
    }
    block_reduce(v, 0u, rows[blockIdx.x].v);
}

// One workgroup: k_pattern_blocks' nrows rows into *out (pinned host memory
// mapped for the device).
__global__ __launch_bounds__(kThreads) void k_pattern_final(const AnalyzeRow* rows, int nrows, AnalyzeStats* out) {
    __shared__ unsigned long long tot[10];
    reduce_rows<10>(rows, nrows, 0u, tot);
    if (threadIdx.x == 0) {
        AnalyzeStats z = {};
        for (int k = 0; k < 4; ++k) {
            z.par_sum[k] = tot[k];
            z.par_cnt[k] = (uint32_t)tot[4 + k];
        }
        z.vote_rgbg = (uint32_t)tot[8];
        z.vote_grgb = (uint32_t)tot[9];
        *out = z;
    }
}

}  // namespace

hipError_t launch_analyze_stats(const uint8_t* d, uint64_t n, uint64_t n1, uint64_t n2, const AnalyzeParams& p,
                                bool scan, AnalyzeRow* rows, AnalyzeStats* out, hipStream_t st) {
    uint64_t blocks = kAnalyzeMinBlocks;
    if (scan) {
        const uint64_t want = (n / (16u * kScanU) + kThreads - 1) / kThreads;  // one step of kScanU chunks per thread
        blocks = want < blocks ? blocks : want > (uint64_t)kAnalyzeMaxBlocks ? kAnalyzeMaxBlocks : want;
    }
    hipLaunchKernelGGL(k_analyze, dim3((unsigned)blocks), dim3(kThreads), 0, st, d, n, n1, n2, p, scan ? 1 : 0, rows);
    hipLaunchKernelGGL(k_analyze_final, dim3(1), dim3(kThreads), 0, st, (const AnalyzeRow*)rows, (int)blocks,
                       scan ? 1 : 0, out);
    return hipGetLastError();
}

hipError_t launch_guess_pattern_stats(const FrameParams& f, int samples, int step, int gs, AnalyzeRow* rows,
                                      AnalyzeStats* out, hipStream_t st) {
    const int nsrows = (samples + kThreads - 1) / kThreads;  // samples <= 20000: at most 79 rows
    if (nsrows < 1 || nsrows > kPatternSampleRows) return hipErrorInvalidValue;
    AnalyzeRow* srows = rows;
    AnalyzeRow* brows = rows + kPatternSampleRows;
    hipLaunchKernelGGL(k_pattern_sample, dim3(nsrows), dim3(kThreads), 0, st, f, samples, srows);
    const int nbx = (f.w - 1 + step - 1) / step, nby = (f.h - 1 + step - 1) / step;  // x < w-1, y < h-1
    const long long nblocks = (long long)nbx * nby;
    const int ngx = (f.w + gs - 1) / gs, ngy = (f.h + gs - 1) / gs;
    const long long npos = (long long)ngx * ngy;
    if (nblocks > 0x7FFFFFFFll || npos > 0x7FFFFFFFll) return hipErrorInvalidValue;
    const long long wv = (nblocks + kThreads - 1) / kThreads, wp = (npos + kThreads - 1) / kThreads;
    const int nvote = (int)(wv > kPatternVoteBlocks ? kPatternVoteBlocks : wv);
    const int npar = (int)(wp < 1 ? 1 : wp > kPatternParityBlocks ? kPatternParityBlocks : wp);
    hipLaunchKernelGGL(k_pattern_blocks, dim3(nvote + npar), dim3(kThreads), 0, st, f, step, nbx, (int)nblocks, gs,
                       ngx, (int)npos, nvote, (const AnalyzeRow*)srows, nsrows, brows);
    hipLaunchKernelGGL(k_pattern_final, dim3(1), dim3(kThreads), 0, st, (const AnalyzeRow*)brows, nvote + npar, out);
    return hipGetLastError();
}

}  // namespace imop_hip
