// spectral_group.hip -- FP32 grouped row and update launches of a batch (every size class in one
// launch), in a unit of their own, apart from the per-class kernels.
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_rows_fused_group<float>(const BatchGroupArgs<float>&, int, size_t, hipStream_t);
template hipError_t launch_pencil_update_group<float>(const BatchGroupArgs<float>&, int, size_t, hipStream_t);

}  // namespace gpf
