// spectral_ext_inv64.hip -- FP64 inverse pencil transforms of the extended lengths.
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_pencils_dir_set<double, 1, 1>(const PencilArgs<double>&, int, hipStream_t);

}  // namespace gpf
