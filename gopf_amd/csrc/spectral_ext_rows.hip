// spectral_ext_rows.hip -- FP32 row passes and the update pass of the extended lengths.
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_pencil_update_set<float, 1>(const PencilUpdateArgs<float>&, int, hipStream_t);
template hipError_t launch_rows_r2c_set<float, 1>(const RowR2CArgs<float>&, hipStream_t);
template hipError_t launch_rows_c2r_set<float, 1>(const RowC2RArgs<float>&, hipStream_t);
template hipError_t launch_rows_sync_set<float, 1>(const RowSyncArgs<float>&, hipStream_t);
template hipError_t launch_rows_fused_set<float, 1>(const RowFusedArgs<float>&, hipStream_t);

}  // namespace gpf
