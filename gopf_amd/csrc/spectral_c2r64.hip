// spectral_c2r64.hip -- FP64 instantiations of the half-spectrum-to-real row transforms k_rows_c2r (the core lengths).
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_rows_c2r<double>(const RowC2RArgs<double>&, hipStream_t);

}  // namespace gpf
