// spectral_pencils64.hip -- FP64 instantiations of the Bluestein pencils k_bluestein (convolution lengths
// 128 to 2048); the plain pencils are in spectral_pencils_fwd64.hip and spectral_pencils_inv64.hip.
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_bluestein<double>(const BluesteinArgs<double>&, int, hipStream_t);

}  // namespace gpf
