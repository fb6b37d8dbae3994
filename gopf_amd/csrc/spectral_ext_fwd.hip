// spectral_ext_fwd.hip -- FP32 forward pencil transforms of the extended lengths (2048, radix 5
// and 7), apart from the core lengths' units so that the hot passes' code objects stay small.
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_pencils_dir_set<float, -1, 1>(const PencilArgs<float>&, int, hipStream_t);

}  // namespace gpf
