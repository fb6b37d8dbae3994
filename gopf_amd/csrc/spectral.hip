// spectral.hip -- host helpers of the spectral kernels (supported lengths, grid sizes) and the
// instantiations of the row transforms k_rows_r2c, k_rows_c2r and k_rows_sync.  The other
// kernel families are instantiated in spectral_pencils.hip, spectral_update.hip and
// spectral_rows.hip; the kernels themselves are in spectral_kernels.h.
#include "spectral_kernels.h"

namespace gpf {

bool fft_size_fast(int n) {
    switch (n) {
        case 2: case 4: case 8: case 16: case 32: case 64: case 128: case 256: case 512: case 1024:
        case 12: case 24: case 48: case 96: case 192: case 384: case 768:
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

#define GPF_INST(T)                                                                  \
    template hipError_t launch_rows_r2c<T>(const RowR2CArgs<T>&, hipStream_t);     \
    template hipError_t launch_rows_c2r<T>(const RowC2RArgs<T>&, hipStream_t);     \
    template hipError_t launch_rows_sync<T>(const RowSyncArgs<T>&, hipStream_t);
GPF_INST(float)
GPF_INST(double)
#undef GPF_INST

}  // namespace gpf
