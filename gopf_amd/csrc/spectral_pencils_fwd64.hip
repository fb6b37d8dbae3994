// spectral_pencils_fwd64.hip -- FP64 instantiations of the forward pencil transforms k_pencils (the core
// lengths, every schedule); a unit of its own so that the families compile in parallel.
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_pencils_dir<double, -1>(const PencilArgs<double>&, int, hipStream_t);

}  // namespace gpf
