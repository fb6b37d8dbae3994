// spectral_ext_inv.hip -- FP32 inverse pencil transforms of the extended lengths.
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_pencils_dir_set<float, 1, 1>(const PencilArgs<float>&, int, hipStream_t);

}  // namespace gpf
