// blockred.h -- block-wide FP64 sums and compensated accumulation, shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace gpf {

// Block-wide FP64 sum, returned to every thread; red must hold blockDim.x doubles.  Blocks of
// whole wavefronts sum each wavefront with cross-lane exchanges and then one value per
// wavefront through LDS (two barriers); others use a tree in LDS.  Every thread adds the
// same values in the same order, so the result is deterministic.  A block of one wavefront
// needs no LDS round: after the exchanges every lane holds the same sum (each pair of lanes
// adds the same two operands at every stage).
__device__ __forceinline__ double block_sum(double v, double* red) {
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    if ((nt & 63) == 0) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (nt == 64) return v;
        if ((tid & 63) == 0) red[tid >> 6] = v;
        __syncthreads();
        double r = 0.0;
        for (int w = 0; w < (nt >> 6); ++w) r += red[w];
        __syncthreads();
        return r;
    }
    red[tid] = v;
    __syncthreads();
    int p2 = 1;
    while (p2 < nt) p2 <<= 1;
    for (int s = p2 / 2; s > 0; s >>= 1) {
        if (tid < s && tid + s < nt) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// Neumaier compensated accumulation
struct KSum {
    double s = 0.0, c = 0.0;
    __device__ void add(double v) {
        double t = s + v;
        c += (fabs(s) >= fabs(v)) ? ((s - t) + v) : ((v - t) + s);
        s = t;
    }
    __device__ double get() const { return s + c; }
};

// Sum of n partial sums by the block's first wavefront (blockDim.x >= 64): lane k adds the
// partials k, k + 64, ... with compensation, then the lanes are combined by cross-lane
// exchanges.  The order depends only on n, not on the block size, so every kernel that
// forms a sum of the same partials gets the same value.  Valid in lanes 0 .. 63.
// partials_load / partials_reduce split it for n <= kPartialsFast: every lane's loads are
// issued first (one round trip to memory for several sums), then added in the same order.
constexpr int kPartialsFast = 256;
struct Partials {
    double v[kPartialsFast / 64];
};
__device__ __forceinline__ Partials partials_load(const double* p, long n) {
    Partials r;
#pragma unroll
    for (int j = 0; j < kPartialsFast / 64; ++j) {
        const long i = threadIdx.x + 64L * j;
        r.v[j] = (threadIdx.x < 64 && i < n) ? p[i] : 0.0;
    }
    return r;
}
__device__ __forceinline__ double partials_reduce(const Partials& r, long n) {
    KSum k;
#pragma unroll
    for (int j = 0; j < kPartialsFast / 64; ++j)
        if (threadIdx.x < 64 && threadIdx.x + 64L * j < n) k.add(r.v[j]);
    double v = k.get();
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double partials_sum(const double* p, long n) {
    KSum k;
    if (threadIdx.x < 64)
        for (long i = threadIdx.x; i < n; i += 64) k.add(p[i]);
    double v = k.get();
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One slot of the energy series at the slot a device counter names (replayable from a
// graph): series[2 i] = sum bulk + ceg sum eg, series[2 i + 1] = sum mass, i = *counter,
// then *counter = i + 1.  Called by a whole block of at least 64 threads.  With at most
// kPartialsFast partials per sum, the three sums' loads and the counter's are issued together.
__device__ __forceinline__ void series_slot(const double* bulk, const double* mass, long nb, const double* eg, long ne,
                                            double ceg, double* series, long* counter) {
    double b, m, e;
    long i = 0;
    if (nb <= kPartialsFast && ne <= kPartialsFast) {
        const Partials pb = partials_load(bulk, nb);
        const Partials pm = partials_load(mass, nb);
        const Partials pe = partials_load(eg, ne);
        if (threadIdx.x == 0) i = *counter;
        b = partials_reduce(pb, nb);
        m = partials_reduce(pm, nb);
        e = partials_reduce(pe, ne);
    } else {
        b = partials_sum(bulk, nb);
        m = partials_sum(mass, nb);
        e = partials_sum(eg, ne);
        if (threadIdx.x == 0) i = *counter;
    }
    if (threadIdx.x == 0) {
        series[2 * i] = b + ceg * e;
        series[2 * i + 1] = m;
        *counter = i + 1;
    }
}

}  // namespace gpf
