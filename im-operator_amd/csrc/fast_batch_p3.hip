// Batched HQLI / MG kernels for Bayer pattern 3 (fast_batch_impl.hpp).
#include "fast_batch_impl.hpp"

namespace imop_hip {
hipError_t launch_batch_pattern3(const BatchArgs& a, int alg, int fmt, bool odd, int rows, bool persist,
                                  int blocks, hipStream_t st) {
    return launch_batch_p<3>(a, alg, fmt, odd, rows, persist, blocks, st);
}
}  // namespace imop_hip
