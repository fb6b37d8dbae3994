// spectral_update.hip -- FP32 instantiations of k_pencil_update (the core lengths, with and without the monitor).
#include "spectral_kernels.h"

namespace gpf {

#define GPF_INST(T) \
    template hipError_t launch_pencil_update<T>(const PencilUpdateArgs<T>&, int, hipStream_t);
GPF_INST(float)
#undef GPF_INST

}  // namespace gpf
