// spectral_pencils.hip -- FP32 instantiations of the Bluestein pencils k_bluestein (convolution lengths
// 128 to 2048); the plain pencils are in spectral_pencils_fwd.hip and spectral_pencils_inv.hip.
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_bluestein<float>(const BluesteinArgs<float>&, int, hipStream_t);

}  // namespace gpf
