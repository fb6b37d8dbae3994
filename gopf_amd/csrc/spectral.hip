// spectral.hip -- host helpers of the spectral kernels (supported lengths, grid sizes, the
// direction switch of the pencil transforms).  The kernel families are instantiated in units
// of their own, one per family and precision (spectral_pencils*.hip, spectral_update*.hip,
// spectral_rows*.hip, spectral_r2c*.hip, spectral_c2r*.hip, spectral_sync*.hip), so that they
// compile in parallel; the kernels themselves are in spectral_kernels.h.
#include "spectral_kernels.h"

namespace gpf {

bool fft_size_fast(int n) {
    switch (n) {
#define GPF_C(NN) case NN:
        GPF_FAST_LENGTHS(GPF_C)
#undef GPF_C
            return true;
        default:
            return false;
    }
}

bool group_length(int n) {
    switch (n) {
#define GPF_C(NN) case NN:
        GPF_GROUP_LENGTHS(GPF_C)
#undef GPF_C
            return true;
        default:
            return false;
    }
}

bool fft_size_supported(int n) { return fft_size_fast(n) || (n >= 1 && n <= 64); }

int fft_threads_per_line(int n, int variant) { return threads_per_line(n, variant); }

long rows_blocks(long rows, int TW) { return (rows + TW - 1) / TW; }
long rows_fused_blocks(long rows, int TW) { return ((rows + 1) / 2 + TW - 1) / TW; }

long pencil_blocks(const Seg* seg, int nseg, int TW, int* tiles0) {
    long tot = 0;
    for (int s = 0; s < nseg; ++s) {
        long t = (long)seg[s].na * ((seg[s].nb + TW - 1) / TW);
        if (s == 0 && tiles0) *tiles0 = (int)t;
        tot += t;
    }
    return tot;
}

// instantiated in spectral_pencils_{fwd,inv}{,64}.hip, not here
extern template hipError_t launch_pencils_dir<float, -1>(const PencilArgs<float>&, int, hipStream_t);
extern template hipError_t launch_pencils_dir<float, +1>(const PencilArgs<float>&, int, hipStream_t);
extern template hipError_t launch_pencils_dir<double, -1>(const PencilArgs<double>&, int, hipStream_t);
extern template hipError_t launch_pencils_dir<double, +1>(const PencilArgs<double>&, int, hipStream_t);

template <typename T>
hipError_t launch_pencils(const PencilArgs<T>& a, int n, int dir, hipStream_t s) {
    return dir < 0 ? launch_pencils_dir<T, -1>(a, n, s) : launch_pencils_dir<T, +1>(a, n, s);
}
template hipError_t launch_pencils<float>(const PencilArgs<float>&, int, int, hipStream_t);
template hipError_t launch_pencils<double>(const PencilArgs<double>&, int, int, hipStream_t);


}  // namespace gpf
