// reduce.h -- FP64 reductions of a real device array (mass, sum of squares, min, max and
// the count of non-finite values), the device replacement of the host scan a reference
// driver runs over m.Fields[i].Data to stop an unstable run
// (examples/solverStability/main.go:58-67).
#pragma once

#include <hip/hip_runtime.h>

namespace gpf {

// result layout, in doubles
enum ReduceSlot : int {
    kRedSum = 0,        // sum of the finite values (Neumaier-compensated)
    kRedSumSq = 1,      // sum of their squares (compensated)
    kRedMin = 2,        // minimum finite value (+inf when there is none)
    kRedMax = 3,        // maximum finite value (-inf when there is none)
    kRedNonFinite = 4,  // number of NaN / +-inf values
    kRedOut = 5
};

// Every block reduces a fixed chunk of kReduceChunk consecutive values and writes kRedOut
// partials; one block then combines the partials in block order.  The partition depends on
// n only, so every entry point that reduces the same array gets the same bits.
constexpr long kReduceChunk = 65536;
// partial blocks the no-allocation entry point keeps in its static device array
constexpr long kReduceStaticBlocks = 4096;
constexpr long kReduceStaticMaxN = kReduceChunk * kReduceStaticBlocks;  // 2^28 values

inline long reduce_blocks(long n) { return n <= 0 ? 1 : (n + kReduceChunk - 1) / kReduceChunk; }

// x[i] + (mean ? *mean : 0) for i < n, T = float or double; partial holds
// kRedOut * reduce_blocks(n) doubles; out (device) receives kRedOut doubles
template <typename T>
hipError_t launch_field_reduce(const T* x, long n, const double* mean, double* partial, double* out, hipStream_t s);

// The same reduction with its partials in the library's static device array of the current
// device (no allocation); n <= kReduceStaticMaxN.  out is device memory of kRedOut doubles,
// written on stream s without a host synchronisation.  Calls that overlap on one device
// share the array and race.
template <typename T>
void field_reduce_static(const T* x, long n, double* out, hipStream_t s);

}  // namespace gpf
