// spectral_persist64.hip -- the FP64 instantiations of the persistent fused 2D step
// (spectral_persist.h), a unit of its own so that the two precisions compile in parallel.
#include "spectral_persist.h"

namespace gpf {

template hipError_t launch_fused_persist<double>(const PersistLaunch<double>&, hipStream_t);
template void persist_geometry<double>(PersistLaunch<double>&, int);
template hipError_t launch_update_mult_table<double>(const PencilUpdateArgs<double>&, hipStream_t);

}  // namespace gpf
