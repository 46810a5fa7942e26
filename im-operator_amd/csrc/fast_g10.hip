// fast_g10.hip -- instantiations of the register-resident kernel for one
// raw layout (split per layout so the kernels compile in parallel).
#include "fast_impl.hpp"

namespace imop_hip {

hipError_t launch_fast_g10(const FastParams& f, int mode, hipStream_t st) { return launch_f<FMT_G10>(f, mode, st); }

}  // namespace imop_hip
