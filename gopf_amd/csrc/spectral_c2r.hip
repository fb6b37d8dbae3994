// spectral_c2r.hip -- FP32 instantiations of the half-spectrum-to-real row transforms k_rows_c2r (the core lengths).
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_rows_c2r<float>(const RowC2RArgs<float>&, hipStream_t);

}  // namespace gpf
