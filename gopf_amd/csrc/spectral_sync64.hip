// spectral_sync64.hip -- FP64 instantiations of the derived-field row pass k_rows_sync (the core lengths).
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_rows_sync<double>(const RowSyncArgs<double>&, hipStream_t);

}  // namespace gpf
