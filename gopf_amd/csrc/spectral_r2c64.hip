// spectral_r2c64.hip -- FP64 instantiations of the real-to-half-spectrum row transforms k_rows_r2c (the core lengths).
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_rows_r2c<double>(const RowR2CArgs<double>&, hipStream_t);

}  // namespace gpf
