// reduce.hip -- FP64 field reductions (sum, sum of squares, min, max, non-finite count).
//
// Two block-level passes: k_reduce_partial reduces one fixed chunk of the array per block
// into kRedOut partials, k_reduce_final combines the partials in block order.  Sums are
// Neumaier-compensated FP64 whatever the array's type; the order of every addition is
// fixed by n, so results are reproducible bit for bit.
//
// The no-allocation path keeps the partials in `scratch`, a static device array; each
// device holds its own copy, whose address is looked up once per device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "common.h"
#include "gpf.h"
#include "reduce.h"

namespace gpf {

// The no-allocation path's partials: kReduceStaticBlocks blocks of kRedOut doubles
// (160 KiB per device, held until the process ends).  In host code this name is not a
// device address; only scratch_ptr()'s result is passed to kernels.
static __device__ __align__(256) double scratch[kReduceStaticBlocks * kRedOut];

namespace {

constexpr int kRedThreads = 256;

struct RSum {
    double s = 0.0, c = 0.0;
    __device__ void add(double v) {
        const double t = s + v;
        c += (fabs(s) >= fabs(v)) ? ((s - t) + v) : ((v - t) + s);
        s = t;
    }
    __device__ void merge(const RSum& o) {
        add(o.s);
        c += o.c;
    }
    __device__ double get() const { return s + c; }
};

struct RAcc {
    RSum sum, sq;
    double mn = INFINITY, mx = -INFINITY, bad = 0.0;
    __device__ void merge(const RAcc& o) {
        sum.merge(o.sum);
        sq.merge(o.sq);
        mn = fmin(mn, o.mn);
        mx = fmax(mx, o.mx);
        bad += o.bad;
    }
};

// tree over the block in a fixed order; the result is valid in thread 0
__device__ RAcc block_merge(RAcc a) {
    __shared__ double sh[7][kRedThreads];
    const int t = threadIdx.x;
    sh[0][t] = a.sum.s; sh[1][t] = a.sum.c; sh[2][t] = a.sq.s; sh[3][t] = a.sq.c;
    sh[4][t] = a.mn; sh[5][t] = a.mx; sh[6][t] = a.bad;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (t < w) {
            RAcc b;
            b.sum.s = sh[0][t + w]; b.sum.c = sh[1][t + w]; b.sq.s = sh[2][t + w]; b.sq.c = sh[3][t + w];
            b.mn = sh[4][t + w]; b.mx = sh[5][t + w]; b.bad = sh[6][t + w];
            a.merge(b);
            sh[0][t] = a.sum.s; sh[1][t] = a.sum.c; sh[2][t] = a.sq.s; sh[3][t] = a.sq.c;
            sh[4][t] = a.mn; sh[5][t] = a.mx; sh[6][t] = a.bad;
        }
        __syncthreads();
    }
    return a;
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_reduce_partial(const T* __restrict__ x, long n,
                                                                const double* __restrict__ mean,
                                                                double* __restrict__ partial) {
    const long start = (long)blockIdx.x * kReduceChunk;
    const long end = (start + kReduceChunk < n) ? start + kReduceChunk : n;
    const double mu = mean ? *mean : 0.0;
    RAcc a;
    for (long i = start + threadIdx.x; i < end; i += kRedThreads) {
        const double v = (double)x[i] + mu;
        if (isfinite(v)) {
            a.sum.add(v);
            a.sq.add(v * v);
            a.mn = fmin(a.mn, v);
            a.mx = fmax(a.mx, v);
        } else {
            a.bad += 1.0;
        }
    }
    a = block_merge(a);
    if (threadIdx.x == 0) {
        double* p = partial + (long)blockIdx.x * kRedOut;
        p[kRedSum] = a.sum.get();
        p[kRedSumSq] = a.sq.get();
        p[kRedMin] = a.mn;
        p[kRedMax] = a.mx;
        p[kRedNonFinite] = a.bad;
    }
}

__global__ void __launch_bounds__(kRedThreads) k_reduce_final(const double* __restrict__ partial, long nblocks,
                                                              double* __restrict__ out) {
    RAcc a;
    for (long b = threadIdx.x; b < nblocks; b += kRedThreads) {
        const double* p = partial + b * kRedOut;
        a.sum.add(p[kRedSum]);
        a.sq.add(p[kRedSumSq]);
        a.mn = fmin(a.mn, p[kRedMin]);
        a.mx = fmax(a.mx, p[kRedMax]);
        a.bad += p[kRedNonFinite];
    }
    a = block_merge(a);
    if (threadIdx.x == 0) {
        out[kRedSum] = a.sum.get();
        out[kRedSumSq] = a.sq.get();
        out[kRedMin] = a.mn;
        out[kRedMax] = a.mx;
        out[kRedNonFinite] = a.bad;
    }
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device, looked up once per device
double* scratch_ptr() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(GPF_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

template <typename T>
hipError_t launch_field_reduce(const T* x, long n, const double* mean, double* partial, double* out, hipStream_t s) {
    const long nb = reduce_blocks(n);
    hipLaunchKernelGGL(k_reduce_partial<T>, dim3((unsigned)nb), dim3(kRedThreads), 0, s, x, n, mean, partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, s, (const double*)partial, nb, out);
    return hipGetLastError();
}

template <typename T>
void field_reduce_static(const T* x, long n, double* out, hipStream_t s) {
    if (n < 0) throw Error(GPF_EINVAL, "negative length");
    if (n > kReduceStaticMaxN)
        throw Error(GPF_EINVAL, "gpf_reduce_noalloc: n = " + std::to_string(n) + " exceeds its maximum of " +
                                    std::to_string(kReduceStaticMaxN) + " values (use gpf_reduce)");
    double* part = scratch_ptr();
    // every partial slot the final pass reads is written by this call's first pass, so
    // nothing left by an earlier call needs clearing
    HIPCHK(launch_field_reduce<T>(x, n, nullptr, part, out, s));
}

template hipError_t launch_field_reduce<float>(const float*, long, const double*, double*, double*, hipStream_t);
template hipError_t launch_field_reduce<double>(const double*, long, const double*, double*, double*, hipStream_t);
template void field_reduce_static<float>(const float*, long, double*, hipStream_t);
template void field_reduce_static<double>(const double*, long, double*, hipStream_t);

}  // namespace gpf
