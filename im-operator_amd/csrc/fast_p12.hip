// fast_p12.hip -- instantiations of the register-resident kernel for one
// raw layout (split per layout so the kernels compile in parallel).
#include "fast_impl.hpp"

namespace imop_hip {

hipError_t launch_fast_p12(const FastParams& f, int mode, hipStream_t st) { return launch_f<FMT_P12>(f, mode, st); }

}  // namespace imop_hip
