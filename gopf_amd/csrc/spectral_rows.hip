// spectral_rows.hip -- FP32 instantiations of k_rows_fused (every length, schedule, with and without the monitor).
#include "spectral_kernels.h"

namespace gpf {

#define GPF_INST(T) \
    template hipError_t launch_rows_fused<T>(const RowFusedArgs<T>&, hipStream_t); \
    template hipError_t launch_rows_fused_group<T>(const BatchGroupArgs<T>&, int, size_t, hipStream_t);
GPF_INST(float)
#undef GPF_INST

}  // namespace gpf
