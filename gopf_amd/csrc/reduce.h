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

inline long reduce_blocks(long n) { return n <= 0 ? 1 : (n + kReduceChunk - 1) / kReduceChunk; }

// x[i] + (mean ? *mean : 0) for i < n, T = float or double; partial holds
// kRedOut * reduce_blocks(n) doubles; out (device) receives kRedOut doubles
template <typename T>
hipError_t launch_field_reduce(const T* x, long n, const double* mean, double* partial, double* out, hipStream_t s);

}  // namespace gpf
