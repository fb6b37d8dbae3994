// spectral_rows.hip -- FP32 instantiations of k_rows_fused (the core lengths, every schedule, with and without the monitor).
#include "spectral_kernels.h"

namespace gpf {

#define GPF_INST(T) \
    template hipError_t launch_rows_fused<T>(const RowFusedArgs<T>&, hipStream_t);
GPF_INST(float)
#undef GPF_INST

}  // namespace gpf
