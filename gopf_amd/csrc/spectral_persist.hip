// spectral_persist.hip -- the FP32 instantiations of the persistent fused 2D step
// (spectral_persist.h); the FP64 ones are in spectral_persist64.hip.
#include "spectral_persist.h"

namespace gpf {

bool persist_side(int n) { return n == 32 || n == 64 || n == 128 || n == 256; }

template hipError_t launch_fused_persist<float>(const PersistLaunch<float>&, hipStream_t);
template void persist_geometry<float>(PersistLaunch<float>&, int);
template hipError_t launch_update_mult_table<float>(const PencilUpdateArgs<float>&, hipStream_t);

}  // namespace gpf
