// spectral_rows64.hip -- the FP64 instantiations of spectral_rows.hip (a unit of its own so that the two
// precisions compile in parallel).
#include "spectral_kernels.h"

namespace gpf {

#define GPF_INST(T) \
    template hipError_t launch_rows_fused<T>(const RowFusedArgs<T>&, hipStream_t);
GPF_INST(double)
#undef GPF_INST

}  // namespace gpf
