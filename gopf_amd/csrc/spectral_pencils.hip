// spectral_pencils.hip -- instantiations of k_pencils (every length, direction, schedule, lines per thread).
#include "spectral_kernels.h"

namespace gpf {

#define GPF_INST(T) \
    template hipError_t launch_pencils<T>(const PencilArgs<T>&, int, int, hipStream_t);
GPF_INST(float)
GPF_INST(double)
#undef GPF_INST

}  // namespace gpf
