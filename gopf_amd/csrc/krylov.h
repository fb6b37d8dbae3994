// krylov.h -- FP64 vector kernels of the device Newton-Krylov solve behind the
// ImplicitEuler stepper: block dot products against a set of basis vectors, linear
// combinations, normalisation by a device-resident norm and max-norm reductions.  All sums
// are formed in a fixed order (per-thread, then per-block tree, then one block over the
// block partials), so repeated solves give identical results.
#pragma once

#include <hip/hip_runtime.h>

namespace gpf {

// longest Krylov basis of one GMRES cycle (restart length)
constexpr int kMaxKrylov = 64;
constexpr int kKrylovThreads = 256;

// number of blocks (and partial sums per vector) the reductions below use for length n
long krylov_blocks(long n);

// partial[i * nblocks + b] = sum over block b's elements e of V[i * ld + e] * w[e],
// for i < nv; nblocks = krylov_blocks(n)
hipError_t launch_multidot(const double* V, long ld, int nv, const double* w, long n, double* partial,
                           hipStream_t s);
// out[i] = sum_b partial[i * nblocks + b], i < nv
hipError_t launch_reduce_rows(const double* partial, int nv, long nblocks, double* out, hipStream_t s);
// w += sign * sum_{i < nv} h[i] V[i * ld + .]   (h in device memory)
hipError_t launch_combine(const double* V, long ld, int nv, const double* h, double sign, double* w, long n,
                          hipStream_t s);
// out = a x + b y; a null x reads as 0
hipError_t launch_axpby(double a, const double* x, double b, const double* y, double* out, long n,
                        hipStream_t s);
// out = x / sqrt(*nrm2), and 0 where *nrm2 is not positive (nrm2 in device memory)
hipError_t launch_normalize(const double* x, const double* nrm2, double* out, long n, hipStream_t s);
// out[0] = max_e |x[e]| (partial: krylov_blocks(n) doubles of scratch)
hipError_t launch_absmax(const double* x, long n, double* partial, double* out, hipStream_t s);

}  // namespace gpf
