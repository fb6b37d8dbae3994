// spectral_stage.hip -- the fused RK4 stage pencils (k_pencils_stage): each field's stage
// update (rk4.go:35-127, the arithmetic of k_rk4_stage_tab) as the values of its first inverse
// pencil pass, for the core lengths.  FP64 only: at 192^3 the FP64 step went from 1.794 to
// 1.648 ms, while the FP32 pass (100 us for two fields against 89 us for the separate stage
// and pencil passes, its operand loads serialised at occupancy 3) was slower.
#include "kspace.h"
#include "spectral_kernels.h"

namespace gpf {

namespace {

// the elements of one line of field f at positions t + TT e, e in [e0, e0 + C): every load
// first, then the stage arithmetic (in k_rk4_stage_tab's order), FINAL / INIT / kn stored,
// the corrected values into v
template <typename T, int E, int C, bool NEEDL>
__device__ __forceinline__ void stage_chunk(const PencilStageArgs<T>& a, const StageLine<T>& f, cpx<T>* v, int e0,
                                            long segoff, long base, long es, int t, int TT, bool valid) {
    long gi[C];
    T sf[C];
    cpx<T> o1[C], o2[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (e0 + c >= E) break;
        gi[c] = segoff + base + (long)(t + TT * (e0 + c)) * es;
        sf[c] = f.stab[gi[c]];
        o1[c] = (a.stage == 0) ? f.k[gi[c]] : f.final[gi[c]];
        o2[c] = (a.stage == 1 || a.stage == 2) ? f.init[gi[c]] : mk<T>(T(0), T(0));
    }
    double L[C], rx[C], ry[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        L[c] = rx[c] = ry[c] = 0.0;
        if constexpr (NEEDL) {
            if (e0 + c < E) {
                int kx = 0, ky = 0, kz = 0;
                kcoords(a.g, gi[c], a.nmain, kx, ky, kz);
                double fr[3];
                ref_freq(a.g, kx, ky, kz, fr);
                L[c] = -(4.0 * 3.14159265358979323846 * 3.14159265358979323846) *
                       (fr[0] * fr[0] + fr[1] * fr[1] + fr[2] * fr[2]);
            }
        }
    }
    // the terms in their order, each one's bricks for the chunk's elements loaded together
    for (int q = 0; q < f.nterm; ++q) {
        const cpx<T>* bp = (q == 0) ? f.brick[0] : (q == 1) ? f.brick[1] : (q == 2) ? f.brick[2] : f.brick[3];
        const double cq = (q == 0) ? f.coef[0] : (q == 1) ? f.coef[1] : (q == 2) ? f.coef[2] : f.coef[3];
        const int pq = (q == 0) ? f.pw[0] : (q == 1) ? f.pw[1] : (q == 2) ? f.pw[2] : f.pw[3];
        const double* dq = (q == 0) ? f.dmul[0] : (q == 1) ? f.dmul[1] : (q == 2) ? f.dmul[2] : f.dmul[3];
        const double dm = dq ? dq[0] : 1.0;
        cpx<T> b[C];
#pragma unroll
        for (int c = 0; c < C; ++c) b[c] = (bp && e0 + c < E) ? bp[gi[c]] : mk<T>(T(0), T(0));
#pragma unroll
        for (int c = 0; c < C; ++c) {
            double me = NEEDL ? cq * ipow(L[c], pq) : cq;
            if (dq) me *= dm;
            if (!bp) {
                rx[c] += me;
                continue;
            }
            rx[c] += me * (double)b[c].x;
            ry[c] += me * (double)b[c].y;
        }
    }
    const double bpf = a.fp * a.dt, bc = a.fc * a.dt;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (e0 + c >= E) break;
        // the right-hand side rounded to T, as RK4.Step stores it between its passes
        const double qx = (double)(T)rx[c], qy = (double)(T)ry[c];
        const double s = (double)sf[c];
        const long i = gi[c];
        cpx<T> kn;
        if (a.stage == 0) {
            const cpx<T> k = o1[c];
            if (valid) {
                f.init[i] = k;
                f.final[i] = mk<T>((T)((double)k.x + bpf * qx), (T)((double)k.y + bpf * qy));
            }
            kn = mk<T>((T)(s * ((double)k.x + bc * qx)), (T)(s * ((double)k.y + bc * qy)));
        } else if (a.stage < 3) {
            const cpx<T> fin = o1[c], ini = o2[c];
            if (valid) f.final[i] = mk<T>((T)((double)fin.x + bpf * qx), (T)((double)fin.y + bpf * qy));
            kn = mk<T>((T)(s * ((double)ini.x + bc * qx)), (T)(s * ((double)ini.y + bc * qy)));
        } else {
            const cpx<T> fin = o1[c];
            // final is stored in T before the division (rk4.go:57-67)
            const T fx = (T)((double)fin.x + bpf * qx), fy = (T)((double)fin.y + bpf * qy);
            kn = mk<T>((T)(s * (double)fx), (T)(s * (double)fy));
        }
        if (valid && a.store_kn) f.kn[i] = kn;
        v[e0 + c] = valid ? kn : mk<T>(T(0), T(0));
    }
    // keep the next chunk's loads after this chunk's arithmetic (bounded register use)
    asm volatile("" ::: "memory");
}

template <typename T, int N, int V, bool NEEDL>
__global__ void __launch_bounds__(kMaxBlock) k_pencils_stage(PencilStageArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    constexpr int E = LineShape<N, V>::E;
    constexpr int C = 2;  // elements whose loads are in flight together
    const int n = (N > 0) ? N : a.p.n;
    const int TT = (N > 0) ? LineShape<N, V>::T : n;
    const TileInfo ti = decode_tile(a.p, blockIdx.x, gridDim.x);
    const LineMap lm = line_map(ti.es == 1, a.p.TW, TT, n);
    const bool valid = lm.l < ti.nvalid;
    // a line past the end of the tile computes on line 0 of the tile and stores nothing
    const long base = ti.base + (long)(valid ? lm.l : 0) * ti.sb;
    const long segoff = ti.s ? a.nmain : 0;
    // one field per block row (blockIdx.y): several fields in one block would keep each
    // field's operands live across the others' transforms
    const int q = blockIdx.y;
    const StageLine<T>& f = (q == 0) ? a.f[0] : (q == 1) ? a.f[1] : (q == 2) ? a.f[2] : a.f[3];
    cpx<T>* o0 = pick(a.p.arr[0][0], a.p.arr[0][1], ti.s);
    cpx<T>* o1 = pick(a.p.arr[1][0], a.p.arr[1][1], ti.s);
    cpx<T>* o2 = pick(a.p.arr[2][0], a.p.arr[2][1], ti.s);
    cpx<T>* o3 = pick(a.p.arr[3][0], a.p.arr[3][1], ti.s);
    cpx<T>* out = (q == 0) ? o0 : (q == 1) ? o1 : (q == 2) ? o2 : o3;
    cpx<T> v[E];
#pragma unroll
    for (int e0 = 0; e0 < E; e0 += C) stage_chunk<T, E, C, NEEDL>(a, f, v, e0, segoff, base, ti.es, lm.t, TT, valid);
    line_fft<T, N, +1, V>(v, lds + lm.off, lm.es, n, lm.t, a.p.tw);
    store_line<T, E>(v, out, base, ti.es, lm.t, TT, valid);
}

}  // namespace

bool stage_pencil_length(int n) {
    switch (n) {
#define GPF_C(NN) case NN:
        GPF_CORE_LENGTHS(GPF_C)
#undef GPF_C
        return true;
        default:
            return n >= 1 && n <= 64;
    }
}

template <typename T>
hipError_t launch_pencils_stage(const PencilStageArgs<T>& a0, int n, hipStream_t s) {
    if (!stage_pencil_length(n) || a0.p.narr < 1 || a0.p.narr > kMaxArr) return hipErrorInvalidValue;
    PencilStageArgs<T> a = a0;
    a.p.n = n;
    a.p.lpt = 1;
    const int var = a.p.variant;
    const long blocks = pencil_blocks(a.p.seg, a.p.nseg, a.p.TW, &a.p.tiles0);
    const int thr = a.p.TW * threads_per_line(n, var);
    const size_t smem = lds_elems(n, a.p.TW) * sizeof(cpx<T>);
    return dispatch_size<0>(n, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        auto go = [&](auto VC) -> hipError_t {
            constexpr int VV = decltype(VC)::value;
            const dim3 grid((unsigned)blocks, (unsigned)a.p.narr);
            if (a.need_l) return launch_k(k_pencils_stage<T, NN, VV, true>, grid, dim3(thr), smem, s, a);
            return launch_k(k_pencils_stage<T, NN, VV, false>, grid, dim3(thr), smem, s, a);
        };
        if constexpr (plain_variant<NN>() == 0) {
            if (!var) return go(std::integral_constant<int, 0>{});
        }
        return go(std::integral_constant<int, 1>{});
    });
}

template hipError_t launch_pencils_stage<double>(const PencilStageArgs<double>&, int, hipStream_t);

}  // namespace gpf
