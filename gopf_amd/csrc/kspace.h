// kspace.h -- k-space coordinates of the half-spectrum layout and the reference's Freq
// values, shared by the pointwise k-space kernels and the fused stage pencils.
#pragma once

#include "spectral.h"

namespace gpf {

// (kx, ky, kz) of half-spectrum index idx (main part first, then the x-Nyquist plane);
// idx < nk < 2^32, so the divisions are 32-bit multiply-high ones
__device__ __forceinline__ void kcoords(const KGrid& g, long idx, long nmain, int& kx, int& ky, int& kz) {
    const unsigned nxh = (unsigned)(g.nx / 2);
    unsigned r;
    if (idx < nmain) {
        const unsigned u = (unsigned)idx;
        r = g.dxh.div(u);
        kx = (int)(u - r * nxh);
    } else {
        r = (unsigned)(idx - nmain);
        kx = (int)nxh;
    }
    const unsigned z = g.dy.div(r);
    ky = (int)(r - z * (unsigned)g.ny);
    kz = (int)z;
}

__device__ __forceinline__ double wrapf(int k, int n) {
    double f = (double)k / (double)n;
    return (f > 0.5) ? f - 1.0 : f;
}

// reference Freq components (fftWrap.go:57-74): f[0] row (y), f[1] col (x), f[2] depth (z),
// read from the grid's table (the same values wrapf computes)
__device__ __forceinline__ void ref_freq(const KGrid& g, int kx, int ky, int kz, double f[3]) {
    if (g.ftab) {
        f[0] = g.ftab[ky];
        f[1] = g.ftab[g.ny + kx];
        f[2] = g.ftab[g.ny + g.nx / 2 + 1 + kz];
        return;
    }
    f[0] = wrapf(ky, g.ny);
    f[1] = wrapf(kx, g.nx);
    f[2] = (g.ndim == 3) ? wrapf(kz, g.nz) : 0.0;
}

__device__ __forceinline__ double ipow(double x, int p) {
    double v = 1.0;
    for (int q = 0; q < p; ++q) v *= x;
    return v;
}

}  // namespace gpf
