// spectral_pencils_fwd.hip -- FP32 instantiations of the forward pencil transforms k_pencils (the core
// lengths, every schedule, lines per thread); a unit of its own so that the families compile in parallel.
#include "spectral_kernels.h"

namespace gpf {

template hipError_t launch_pencils_dir<float, -1>(const PencilArgs<float>&, int, hipStream_t);

}  // namespace gpf
