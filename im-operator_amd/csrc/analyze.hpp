// analyze.hpp -- device statistics of the raw-data analyzer (analyze_kernels.hip)
// and their launchers; the decisions on them are the host's (imop_analyze.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace imop_hip {

struct AnalyzeParams {
    uint32_t sample[4];  // packed pixels to decode at depths 10, 12, 14, 16
};

struct AnalyzeStats {
    // analyze_data / detect_bit_depth
    uint32_t max16_n1;             // max of the first n1 u16 values
    uint32_t hi_n2, zero_n2;       // of the first n2: high byte non-zero, value 0
    uint32_t max16_n2;             // max of the first n2
    uint32_t min16_all, max16_all; // all n / 2 u16 values
    uint32_t min8_all, max8_all;   // all n bytes
    uint32_t pmin[4], pmax[4];     // the packed samples at depths 10, 12, 14, 16
    uint32_t scanned;              // the whole-buffer ranges were gathered
    // guess_pattern
    uint32_t vote_rgbg, vote_grgb;
    unsigned long long par_sum[4];
    uint32_t par_cnt[4];
};

// One workgroup's partial statistics; a one-workgroup kernel reduces them.
struct AnalyzeRow {
    unsigned long long v[16];
};
constexpr int kAnalyzeMinBlocks = 64;      // k_analyze workgroups without the whole-buffer scan
constexpr int kAnalyzeMaxBlocks = 1024;    // and at most with it
constexpr int kPatternSampleRows = 80;     // ceil(20000 / 256) = 79 sample workgroups
constexpr int kPatternVoteBlocks = 256;
constexpr int kPatternParityBlocks = 1024;
constexpr int kAnalyzeRowCount = 2048;     // rows the caller provides, >= either launcher's need
static_assert(kAnalyzeMaxBlocks <= kAnalyzeRowCount, "rows");
static_assert(kPatternSampleRows + kPatternVoteBlocks + kPatternParityBlocks <= kAnalyzeRowCount, "rows");

// k_analyze, then k_analyze_final: the statistics into *out (device-visible;
// pinned host memory in the library), the whole-buffer ranges only with scan.
hipError_t launch_analyze_stats(const uint8_t* d, uint64_t n, uint64_t n1, uint64_t n2, const AnalyzeParams& p,
                                bool scan, AnalyzeRow* rows, AnalyzeStats* out, hipStream_t st);
// k_pattern_sample, k_pattern_blocks, k_pattern_final: the votes and the
// parity sums of imop::guess_pattern into *out.
hipError_t launch_guess_pattern_stats(const FrameParams& f, int samples, int step, int gs, AnalyzeRow* rows,
                                      AnalyzeStats* out, hipStream_t st);

}  // namespace imop_hip
