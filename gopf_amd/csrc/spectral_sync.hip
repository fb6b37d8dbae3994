// spectral_sync.hip -- FP32 instantiations of the derived-field row pass k_rows_sync (the core lengths).
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_rows_sync<float>(const RowSyncArgs<float>&, hipStream_t);

}  // namespace gpf
