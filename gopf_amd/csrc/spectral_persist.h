// spectral_persist.h -- every step of a run of fused 2D steps in one cooperative launch
// (Euler.Step, pf/euler.go:16-47, in the fused form of k_rows_fused and k_pencil_update).
//
// A small grid's step is two dependent passes of a few microseconds of work each, and a
// launch per pass costs more than the pass.  Here the launch's workgroups loop over the
// steps: each runs its share of the row pass, all meet at a device-wide barrier, run their
// share of the update pass, and meet again.  The passes are the bodies of the per-pass
// kernels (same arithmetic, same partial sums per block of the same layout), with the
// energy-series slot of step s written during step s + 1's row pass (by the participant the
// row pass leaves idle, else by block 0 of that pass).
//
// The barrier is a counter and a generation word in device memory.  Its wait is bounded: a
// workgroup that waits longer than the
// spin limit (or finds another's failure) raises the failure word, sets the caller's mapped
// host flag and returns, and every other workgroup then returns at its next barrier, so the
// launch always ends.  Co-residency of the grid comes from the cooperative launch, which
// refuses a grid the device cannot hold at once.
#pragma once

#include "spectral_kernels.h"

namespace gpf {

// The launch's participating workgroups: with LOCAL, blocks b with b % 8 == 0 only, which the
// device's round-robin dispatch of workgroups over its 8 XCDs places on one XCD (checked at
// the start of the launch from the XCC_ID hardware register); the others return at once.
// One XCD shares one L2, so a barrier hands data over when every wave's stores have reached
// the L2 (s_waitcnt vmcnt(0)) and the waiting blocks' L1 is invalidated (buffer_inv sc1): no
// L2 writeback is needed, and the passes read each other's results from the L2.  Without
// LOCAL the barrier is agent-scope release / acquire fences in every wave.
//
// barrier words: [0] arrivals (monotonic within a launch), [1] generation, [2] failure,
// [3] mask of the XCC ids of the participants
template <bool LOCAL>
__device__ __forceinline__ bool persist_barrier(unsigned* bar, unsigned& gen, unsigned np, unsigned spin_limit,
                                                int* host_fail) {
    if constexpr (LOCAL) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    int failed = 0;
    if (threadIdx.x == 0) {
        const unsigned g = gen;
        if (__hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
            failed = 1;
        } else {
            // the arrivals of barrier g are np (g + 1): the count is never reset, so the
            // generation store is the only write of the last arrival
            const unsigned arrived =
                __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
            if (arrived == np * (g + 1u)) {
                __hip_atomic_store(&bar[1], g + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                unsigned it = 0;
                while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                    if (++it > spin_limit || __hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        failed = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        if (failed) {
            __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(host_fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        gen = g + 1u;
        if constexpr (LOCAL) asm volatile("buffer_inv sc1" ::: "memory");
    }
    failed = __syncthreads_or(failed);
    if constexpr (!LOCAL) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return failed == 0;
}

// the step's arguments are the kernel argument P (2.4 KB, read with scalar loads from the
// constant argument segment: read from global memory, every pass would wait on vector loads
// of its coefficients): the row pass of a launch's first step (no series slot to write), of
// odd steps (partial sums in set 1, slot of set 0) and of even steps after the first (set 0,
// slot of set 1); the update pass of even and odd steps
template <typename T, int N, bool MON, bool LOCAL>
__global__ void __launch_bounds__(kMaxBlock)
    k_fused_persist(const PersistPlan<T> P, int steps, unsigned* bar, unsigned spin_limit, int* host_fail) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr unsigned kStride = LOCAL ? 8u : 1u;
    if (blockIdx.x % kStride) return;
    const int pb = (int)(blockIdx.x / kStride);
    const unsigned np = (gridDim.x + kStride - 1u) / kStride;
    unsigned gen = 0;
    if constexpr (LOCAL) {
        // every participant records its XCC; after a fenced barrier all read the same mask
        // and, unless it names one XCC, return before touching the fields (host flag 2)
        if (threadIdx.x == 0) {
            const unsigned xcc = __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 15u;  // HW_REG_XCC_ID
            __hip_atomic_fetch_or(&bar[3], 1u << xcc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
        }
        if (!persist_barrier<false>(bar, gen, np, spin_limit, host_fail)) return;
        int other = 0;
        if (threadIdx.x == 0) {
            const unsigned mask = __hip_atomic_load(&bar[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            other = (mask & (mask - 1u)) != 0u;
            if (other && pb == 0) __hip_atomic_store(host_fail, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        if (__syncthreads_or(other)) return;
    }
    for (int s = 0; s < steps; ++s) {
        if (P.rows_pass) {
            const RowFusedArgs<T>& ra = P.ra[s == 0 ? 0 : ((s & 1) ? 1 : 2)];
            for (int b = pb; b < P.nblk_rows; b += (int)np) {
                rows_fused_body<T, N, 1, MON, const RowFusedArgs<T>&>(ra, b, smem);
                __syncthreads();
            }
            if (MON && P.slot_series && s > 0 && pb == (int)np - 1)
                series_slot(ra.prev_bulk, ra.prev_mass, ra.prev_nb, ra.prev_eg, ra.prev_ne, ra.prev_ceg, P.slot_series,
                            P.slot_counter);
            if (!persist_barrier<LOCAL>(bar, gen, np, spin_limit, host_fail)) return;
        }
        const PencilUpdateArgs<T>& u = P.u[s & 1];
        for (int b = pb; b < P.nblk_upd; b += (int)np) {
            pencil_update_body<T, N, MON, const PencilUpdateArgs<T>&, true, 1>(u, b, P.nblk_upd, smem);
            __syncthreads();
        }
        if (s + 1 < steps && !persist_barrier<LOCAL>(bar, gen, np, spin_limit, host_fail)) return;
    }
}

// the update pass's multipliers of every k-point (the pass in its table-writing form)
template <typename T, int N>
__global__ void __launch_bounds__(kMaxBlock) k_update_mult_table(PencilUpdateArgs<T> u) {
    extern __shared__ __align__(16) unsigned char smem[];
    pencil_update_body<T, N, false, PencilUpdateArgs<T>, false, 2>(u, blockIdx.x, gridDim.x, smem);
}

template <typename T>
hipError_t launch_update_mult_table(const PencilUpdateArgs<T>& u0, hipStream_t s) {
    PencilUpdateArgs<T> u = u0;
    u.p.lpt = 1;
    const long blocks = pencil_blocks(u.p.seg, u.p.nseg, u.p.TW, &u.p.tiles0);
    const int thr = u.p.TW * threads_per_line(u.p.n, 1);
    auto go = [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        auto kern = k_update_mult_table<T, NN>;
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(thr), 0, s, u);
        return hipGetLastError();
    };
    switch (u.p.n) {
        case 32: return go(std::integral_constant<int, 32>{});
        case 64: return go(std::integral_constant<int, 64>{});
        case 128: return go(std::integral_constant<int, 128>{});
        case 256: return go(std::integral_constant<int, 256>{});
        default: return hipErrorInvalidValue;
    }
}

// threads per block (tw lines of the schedule-1 transform) and dynamic LDS: the larger of the
// row pass's (exchange lines, twiddle copy) and the update pass's, with the monitor's sums
template <typename T>
void persist_geometry(PersistLaunch<T>& L, int tw) {
    L.threads = tw * threads_per_line(L.n, 1);
    L.smem = (lds_elems(L.n, tw) + L.n) * sizeof(cpx<T>) + (L.mon ? (size_t)L.threads * sizeof(double) : 0);
}

template <typename T>
hipError_t launch_fused_persist(const PersistLaunch<T>& L, hipStream_t s) {
    auto go = [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        void* kern = L.local ? (L.mon ? reinterpret_cast<void*>(k_fused_persist<T, NN, true, true>)
                                      : reinterpret_cast<void*>(k_fused_persist<T, NN, false, true>))
                             : (L.mon ? reinterpret_cast<void*>(k_fused_persist<T, NN, true, false>)
                                      : reinterpret_cast<void*>(k_fused_persist<T, NN, false, false>));
        PersistPlan<T> plan = L.plan;
        int steps = L.steps;
        unsigned* bar = L.bar;
        unsigned spin = L.spin_limit;
        int* hf = L.host_fail;
        void* args[] = {&plan, &steps, &bar, &spin, &hf};
        const unsigned grid = (unsigned)L.blocks * (L.local ? 8u : 1u);
        return hipLaunchCooperativeKernel(kern, dim3(grid), dim3(L.threads), args, (unsigned)L.smem, s);
    };
    switch (L.n) {
        case 32: return go(std::integral_constant<int, 32>{});
        case 64: return go(std::integral_constant<int, 64>{});
        case 128: return go(std::integral_constant<int, 128>{});
        case 256: return go(std::integral_constant<int, 256>{});
        default: return hipErrorInvalidValue;
    }
}

}  // namespace gpf
