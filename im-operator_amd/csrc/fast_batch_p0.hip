// Batched HQLI / MG kernels for Bayer pattern 0 (fast_batch_impl.hpp).
#include "fast_batch_impl.hpp"

namespace imop_hip {
hipError_t launch_batch_pattern0(const BatchArgs& a, int alg, int rows, int blocks, hipStream_t st) {
    return launch_batch_p<0>(a, alg, rows, blocks, st);
}
}  // namespace imop_hip
