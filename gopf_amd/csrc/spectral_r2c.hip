// spectral_r2c.hip -- FP32 instantiations of the real-to-half-spectrum row transforms k_rows_r2c (the core lengths).
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_rows_r2c<float>(const RowR2CArgs<float>&, hipStream_t);

}  // namespace gpf
