// spectral_ext_rows64.hip -- FP64 row passes and the update pass of the extended lengths.
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_pencil_update_set<double, 1>(const PencilUpdateArgs<double>&, int, hipStream_t);
template hipError_t launch_rows_r2c_set<double, 1>(const RowR2CArgs<double>&, hipStream_t);
template hipError_t launch_rows_c2r_set<double, 1>(const RowC2RArgs<double>&, hipStream_t);
template hipError_t launch_rows_sync_set<double, 1>(const RowSyncArgs<double>&, hipStream_t);
template hipError_t launch_rows_fused_set<double, 1>(const RowFusedArgs<double>&, hipStream_t);

}  // namespace gpf
