// krylov.hip -- FP64 vector kernels of the device Newton-Krylov solve (ImplicitEuler).
//
// The GMRES driver keeps its basis as nv vectors of length n at a fixed leading dimension
// ld.  Classical Gram-Schmidt needs the nv dot products <V_i, w> in one sweep: k_multidot
// reads w once per group of kGroup basis vectors and each basis vector once, with the
// group's accumulators in registers (the group index is a compile-time constant in the
// unrolled loop), and writes one partial per block and vector; k_reduce_rows folds the
// partials.  Every sum has a fixed order.
#include "krylov.h"

namespace gpf {

namespace {

constexpr int kGroup = 8;
constexpr int kWaves = kKrylovThreads / 64;

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// max that propagates NaN (fmax drops it), so a non-finite residual is never read as small
__device__ __forceinline__ double nanmax(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    return a > b ? a : b;
}

__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o));
    return v;
}

__global__ void __launch_bounds__(kKrylovThreads) k_multidot(const double* __restrict__ V, long ld, int nv,
                                                              const double* __restrict__ w, long n,
                                                              double* __restrict__ partial) {
    const int i0 = blockIdx.y * kGroup;
    const int cnt = min(kGroup, nv - i0);
    double acc[kGroup];
#pragma unroll
    for (int q = 0; q < kGroup; ++q) acc[q] = 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride) {
        const double we = w[e];
#pragma unroll
        for (int q = 0; q < kGroup; ++q)
            if (q < cnt) acc[q] += V[(long)(i0 + q) * ld + e] * we;
    }
    __shared__ double red[kGroup][kWaves];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < kGroup; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == 0) red[q][wv] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
        double t = 0.0;
        for (int k = 0; k < kWaves; ++k) t += red[threadIdx.x][k];
        partial[(long)(i0 + threadIdx.x) * gridDim.x + blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kKrylovThreads) k_reduce_rows(const double* __restrict__ partial, long nblocks,
                                                                 double* __restrict__ out) {
    const double* row = partial + (long)blockIdx.x * nblocks;
    double v = 0.0;
    for (long b = threadIdx.x; b < nblocks; b += blockDim.x) v += row[b];
    __shared__ double red[kWaves];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kWaves; ++k) t += red[k];
        out[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kKrylovThreads) k_combine(const double* __restrict__ V, long ld, int nv,
                                                             const double* __restrict__ h, double sign,
                                                             double* __restrict__ w, long n) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride) {
        double acc = 0.0;
        for (int i = 0; i < nv; ++i) acc += h[i] * V[(long)i * ld + e];
        w[e] += sign * acc;
    }
}

// out may alias x or y
__global__ void __launch_bounds__(kKrylovThreads) k_axpby(double a, const double* x, double b, const double* y,
                                                           double* out, long n) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride)
        out[e] = (x ? a * x[e] : 0.0) + b * y[e];
}

__global__ void __launch_bounds__(kKrylovThreads) k_normalize(const double* __restrict__ x,
                                                               const double* __restrict__ nrm2,
                                                               double* __restrict__ out, long n) {
    const double s2 = nrm2[0];
    const double inv = (s2 > 0.0) ? 1.0 / sqrt(s2) : 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride) out[e] = x[e] * inv;
}

__global__ void __launch_bounds__(kKrylovThreads) k_absmax(const double* __restrict__ x, long n,
                                                            double* __restrict__ partial) {
    double v = 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride) v = nanmax(v, fabs(x[e]));
    __shared__ double red[kWaves];
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kWaves; ++k) t = nanmax(t, red[k]);
        partial[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kKrylovThreads) k_max_rows(const double* __restrict__ partial, long nblocks,
                                                              double* __restrict__ out) {
    double v = 0.0;
    for (long b = threadIdx.x; b < nblocks; b += blockDim.x) v = nanmax(v, partial[b]);
    __shared__ double red[kWaves];
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kWaves; ++k) t = nanmax(t, red[k]);
        out[0] = t;
    }
}

inline unsigned grid_for(long n) {
    long g = (n + kKrylovThreads - 1) / kKrylovThreads;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

long krylov_blocks(long n) {
    long g = (n + kKrylovThreads - 1) / kKrylovThreads;
    if (g > 1024) g = 1024;
    if (g < 1) g = 1;
    return g;
}

hipError_t launch_multidot(const double* V, long ld, int nv, const double* w, long n, double* partial,
                           hipStream_t s) {
    if (nv <= 0) return hipSuccess;
    const unsigned groups = (unsigned)((nv + kGroup - 1) / kGroup);
    hipLaunchKernelGGL(k_multidot, dim3((unsigned)krylov_blocks(n), groups), dim3(kKrylovThreads), 0, s, V, ld, nv,
                       w, n, partial);
    return hipGetLastError();
}

hipError_t launch_reduce_rows(const double* partial, int nv, long nblocks, double* out, hipStream_t s) {
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_rows, dim3((unsigned)nv), dim3(kKrylovThreads), 0, s, partial, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_combine(const double* V, long ld, int nv, const double* h, double sign, double* w, long n,
                          hipStream_t s) {
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_combine, dim3(grid_for(n)), dim3(kKrylovThreads), 0, s, V, ld, nv, h, sign, w, n);
    return hipGetLastError();
}

hipError_t launch_axpby(double a, const double* x, double b, const double* y, double* out, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kKrylovThreads), 0, s, a, x, b, y, out, n);
    return hipGetLastError();
}

hipError_t launch_normalize(const double* x, const double* nrm2, double* out, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_normalize, dim3(grid_for(n)), dim3(kKrylovThreads), 0, s, x, nrm2, out, n);
    return hipGetLastError();
}

hipError_t launch_absmax(const double* x, long n, double* partial, double* out, hipStream_t s) {
    const long nb = krylov_blocks(n);
    hipLaunchKernelGGL(k_absmax, dim3((unsigned)nb), dim3(kKrylovThreads), 0, s, x, n, partial);
    hipLaunchKernelGGL(k_max_rows, dim3(1), dim3(kKrylovThreads), 0, s, partial, nb, out);
    return hipGetLastError();
}

}  // namespace gpf
