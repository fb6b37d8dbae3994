// spectral_kernels.h -- hand-written gfx950 kernels of gopf's spectral time step and their
// launch templates.  Included by the spectral*.hip translation units, each of which
// instantiates one kernel family (the instantiations of every transform length are what
// makes these units slow to compile, so they are compiled in parallel).
//
// Every kernel is one HBM pass: a tile of lines is staged into LDS with
// coalesced loads, transformed in LDS (fft_core.h), optionally combined with
// pointwise / k-space work, and written back with coalesced stores.
//
// Kernels
//   k_pencils        C2C transform of pencils along a strided or contiguous axis
//                    (y / z passes over the half spectrum, and every axis of the
//                    full-complex FourierTransform drop-in, fftWrap.go:26-39)
//   k_pencil_update  forward transform of the explicit source along the last axis ->
//                    semi-implicit update of the state spectrum (euler.go:27-39 with
//                    LaplacianN.Eval diffOp.go:25-29, ApplyModalFilter util.go:125-132)
//                    -> inverse transform of the new state along that axis
//   k_rows_r2c       x-row transforms of pairs of real arrays into half spectra
//   k_rows_c2r       x-row inverse of half spectra into real rows (euler.go:42-45)
//   k_rows_fused     x-row inverse -> pointwise explicit source g(c) (SyncDerivedFields,
//                    model.go:237-241) -> x-row forward of g, two rows per transform;
//                    optional FP64 energy/mass partial sums
//
// The fused step keeps the state as its spectrum S = FFT(c - mean) / N: the reference
// transforms the field forward and back every step (euler.go:20-45), which here would
// add the rounding of two full transforms to the state each step; only the explicit
// source, a small correction, passes through the transforms.
#pragma once

#include <type_traits>

#include "blockred.h"
#include "fft_core.h"
#include "spectral.h"

namespace gpf {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kMaxBlock = kMaxBlockThreads;

// elements per thread and threads per line of the transform of length n
template <int N, int V = 0>
struct LineShape {
    static constexpr int E = (N > 0) ? Sched<(N > 0 ? N : 2), V>::E : 1;
    static constexpr int T = (N > 0) ? Sched<(N > 0 ? N : 2), V>::T : 0;  // 0: runtime (= n)
};

// schedule of the plain passes: the low-register variant once a thread would hold 24+ elements
template <int N>
constexpr int plain_variant() {
    return (N > 0 && LineShape<N, 0>::E >= 24) ? 1 : 0;
}

template <typename T, int N, int DIR, int V = 0>
__device__ __forceinline__ void line_fft(cpx<T>* v, cpx<T>* lds, int es, int n, int t, const cpx<T>* __restrict__ tw) {
    if constexpr (N > 0) {
        fft_canon<T, N, DIR, V>(v, lds, es, t, tw);
    } else {
        dft_canon_rt<T, DIR>(v, lds, es, n, t, tw);
    }
}

// 1/x in FP64: correctly rounded division for the plain FP64 path; hardware
// reciprocal refined by two Newton steps (relative error ~1e-16) for FP32-bulk
template <typename T>
__device__ __forceinline__ double recip(double x) {
    if constexpr (sizeof(T) == 8) {
        return 1.0 / x;
    } else {
        double r = __builtin_amdgcn_rcp(x);
        r = r * (2.0 - x * r);
        r = r * (2.0 - x * r);
        return r;
    }
}

__device__ __forceinline__ int wrapk(int k, int n) { return (2 * k > n) ? k - n : k; }

__device__ __forceinline__ double kpoly_eval(const KPoly& p, double L) {
    double acc = 0.0;
    for (int t = 0; t < p.n; ++t) {
        double v = p.coef[t];
        for (int q = 0; q < p.pow[t]; ++q) v *= L;
        acc += v;
    }
    return acc;
}

__device__ __forceinline__ double vandeven(const double* __restrict__ tab, double x) {
    const double N = 999.0;
    int idx = (int)(x * N);
    if (idx >= 999) return tab[999];
    double dx = 1.0 / N;
    double dy = tab[idx + 1] - tab[idx];
    double x0 = (double)idx * dx;
    return tab[idx] + (x - x0) * dy / dx;
}



// ---------------------------------------------------------------------------
// tile decode shared by the pencil kernels
// ---------------------------------------------------------------------------
struct TileInfo {
    int s;        // segment
    int ia, b0;   // a index and first b of the tile
    long base;    // element offset of line (ia, b0)
    int nvalid;   // valid lines in the tile
    long sb, es;  // line and element strides of the segment
    int ksrc0, ksrc1, ksrc2, kfix0, kfix1, kfix2;
    int sim_src;
};

// The segment is chosen with selects (no dynamic indexing of the argument
// struct, which would copy it to scratch memory).
// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8), and
// each XCD has its own L2.  Neighbouring tiles of a plane share cache lines when a tile
// row is narrower than a line, so consecutive tiles go to workgroups of one XCD: XCD x
// takes the x-th contiguous run of tiles (a bijection of [0, nblk)).
__device__ __forceinline__ int xcd_tile(int b, int nblk) {
    const int q = nblk >> 3, r = nblk & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}

// tile of block b of a launch of nblk blocks (a grouped launch passes a class's own range)
template <typename T>
__device__ __forceinline__ TileInfo decode_tile(const PencilArgs<T>& a, int b, int nblk) {
    TileInfo ti;
    int tile = xcd_tile(b, nblk);
    const bool s1 = tile >= a.tiles0;
    ti.s = s1 ? 1 : 0;
    if (s1) tile -= a.tiles0;
    const int nb = s1 ? a.seg[1].nb : a.seg[0].nb;
    const long sa = s1 ? a.seg[1].sa : a.seg[0].sa;
    ti.sb = s1 ? a.seg[1].sb : a.seg[0].sb;
    ti.es = s1 ? a.seg[1].es : a.seg[0].es;
    ti.ksrc0 = s1 ? a.seg[1].ksrc[0] : a.seg[0].ksrc[0];
    ti.ksrc1 = s1 ? a.seg[1].ksrc[1] : a.seg[0].ksrc[1];
    ti.ksrc2 = s1 ? a.seg[1].ksrc[2] : a.seg[0].ksrc[2];
    ti.kfix0 = s1 ? a.seg[1].kfix[0] : a.seg[0].kfix[0];
    ti.kfix1 = s1 ? a.seg[1].kfix[1] : a.seg[0].kfix[1];
    ti.kfix2 = s1 ? a.seg[1].kfix[2] : a.seg[0].kfix[2];
    ti.sim_src = s1 ? a.seg[1].sim_src : a.seg[0].sim_src;
    const int tl = a.TW * a.lpt;  // lines per tile
    const int tpa = (nb + tl - 1) / tl;
    ti.ia = tile / tpa;
    ti.b0 = (tile - ti.ia * tpa) * tl;
    ti.base = (long)ti.ia * sa + (long)ti.b0 * ti.sb;
    ti.nvalid = min(tl, nb - ti.b0);
    return ti;
}

template <typename P>
__device__ __forceinline__ P pick(P a0, P a1, int s) { return s ? a1 : a0; }

__device__ __forceinline__ int ksel(int src, int fix, int a, int b, int p) {
    return (src == 0) ? a : (src == 1) ? b : (src == 2) ? p : fix;
}

// Thread -> (line l, thread-in-line t).  Contiguous lines (es == 1) put t
// fastest so a wave reads consecutive elements of a line; strided lines put
// l fastest so a wave reads adjacent lines (consecutive addresses).
struct LineMap {
    int l, t;
    int es;          // LDS element stride
    int off;         // LDS offset of this thread's line
};

__device__ __forceinline__ LineMap line_map(bool contiguous, int TW, int TT, int n) {
    LineMap m;
    if (contiguous) {
        m.l = threadIdx.x / TT;
        m.t = threadIdx.x - m.l * TT;
        m.es = 1;
        m.off = m.l * lds_line(n);
    } else {
        m.l = threadIdx.x & (TW - 1);
        m.t = threadIdx.x / TW;
        m.es = TW + 1;
        m.off = m.l;
    }
    return m;
}

__host__ __device__ constexpr size_t lds_elems(int n, int TW) {
    // both layouts fit in lds_line(n) * (TW + 1) elements
    return (size_t)lds_line(n) * (size_t)(TW + 1);
}

template <typename T, int E>
__device__ __forceinline__ void load_line(cpx<T>* v, const cpx<T>* __restrict__ src, long base, long es, int t, int TT,
                                          bool valid) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = valid ? src[base + (long)(t + TT * e) * es] : mk<T>(T(0), T(0));
}

template <typename T, int E>
__device__ __forceinline__ void store_line(const cpx<T>* v, cpx<T>* __restrict__ dst, long base, long es, int t, int TT,
                                           bool valid) {
    if (!valid) return;
#pragma unroll
    for (int e = 0; e < E; ++e) dst[base + (long)(t + TT * e) * es] = v[e];
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// LPT = 2: thread-line l transforms the adjacent lines 2l and 2l + 1 of a tile of 2 TW
// lines, both loaded before either is transformed.  When the lines are adjacent in memory
// (line stride 1) and the pair is 16-byte aligned, FP32 pairs move as one 16-byte vector
// per element, so eight lanes cover a whole 128-byte cache line.
template <typename T, int E>
__device__ __forceinline__ void load_line_pair(cpx<T>* v, cpx<T>* w, const cpx<T>* __restrict__ src, long base, long es,
                                               int t, int TT, bool vec, bool valid0, bool valid1) {
    if constexpr (sizeof(T) == 4) {
        if (vec) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const float4 x = *reinterpret_cast<const float4*>(src + base + (long)(t + TT * e) * es);
                v[e] = mk<T>(x.x, x.y);
                w[e] = mk<T>(x.z, x.w);
            }
            return;
        }
    }
    load_line<T, E>(v, src, base, es, t, TT, valid0);
    load_line<T, E>(w, src, base + 1, es, t, TT, valid1);
}

template <typename T, int E>
__device__ __forceinline__ void store_line_pair(const cpx<T>* v, const cpx<T>* w, cpx<T>* __restrict__ dst, long base,
                                                long es, int t, int TT, bool vec, bool valid0, bool valid1) {
    if constexpr (sizeof(T) == 4) {
        if (vec) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                float4 x;
                x.x = v[e].x;
                x.y = v[e].y;
                x.z = w[e].x;
                x.w = w[e].y;
                *reinterpret_cast<float4*>(dst + base + (long)(t + TT * e) * es) = x;
            }
            return;
        }
    }
    store_line<T, E>(v, dst, base, es, t, TT, valid0);
    store_line<T, E>(w, dst, base + 1, es, t, TT, valid1);
}

template <typename T, int N, int DIR, int V, int LPT>
__global__ void __launch_bounds__(kMaxBlock) k_pencils(PencilArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    constexpr int E = LineShape<N, V>::E;
    const int n = (N > 0) ? N : a.n;
    const int TT = (N > 0) ? LineShape<N, V>::T : n;
    const TileInfo ti = decode_tile(a, blockIdx.x, gridDim.x);
    const LineMap lm = line_map(ti.es == 1, a.TW, TT, n);
#pragma unroll
    for (int q = 0; q < kMaxArr; ++q) {
        if (q < a.narr) {
            cpx<T>* out = pick(a.arr[q][0], a.arr[q][1], ti.s);
            const cpx<T>* s = pick(a.src[q][0], a.src[q][1], ti.s);
            const cpx<T>* in = s ? s : out;
            if constexpr (LPT == 1) {
                const bool valid0 = lm.l < ti.nvalid;
                const long base0 = ti.base + (long)lm.l * ti.sb;
                cpx<T> v[E];
                load_line<T, E>(v, in, base0, ti.es, lm.t, TT, valid0);
                line_fft<T, N, DIR, V>(v, lds + lm.off, lm.es, n, lm.t, a.tw);
                store_line<T, E>(v, out, base0, ti.es, lm.t, TT, valid0);
            } else {
                const int l0 = 2 * lm.l;
                const bool valid0 = l0 < ti.nvalid, valid1 = l0 + 1 < ti.nvalid;
                const long base0 = ti.base + (long)l0 * ti.sb;
                const bool vec = valid1 && ti.sb == 1 && (ti.es & 1) == 0 &&
                                 ((reinterpret_cast<size_t>(in + base0) | reinterpret_cast<size_t>(out + base0)) & 15) == 0;
                cpx<T> v[E], w[E];
                if (ti.sb == 1) {
                    load_line_pair<T, E>(v, w, in, base0, ti.es, lm.t, TT, vec, valid0, valid1);
                } else {
                    load_line<T, E>(v, in, base0, ti.es, lm.t, TT, valid0);
                    load_line<T, E>(w, in, base0 + ti.sb, ti.es, lm.t, TT, valid1);
                }
                line_fft<T, N, DIR, V>(v, lds + lm.off, lm.es, n, lm.t, a.tw);
                line_fft<T, N, DIR, V>(w, lds + lm.off, lm.es, n, lm.t, a.tw);
                if (ti.sb == 1) {
                    store_line_pair<T, E>(v, w, out, base0, ti.es, lm.t, TT, vec, valid0, valid1);
                } else {
                    store_line<T, E>(v, out, base0, ti.es, lm.t, TT, valid0);
                    store_line<T, E>(w, out, base0 + ti.sb, ti.es, lm.t, TT, valid1);
                }
            }
        }
    }
}

// Bluestein pencils: a length-n line (n at run time) is multiplied by the chirp, zero-padded
// to M, transformed forward, multiplied by the chirp filter's spectrum, transformed back,
// and multiplied by the chirp again; positions p >= n of the M-point line are discarded.
template <typename T, int M>
__global__ void __launch_bounds__(kMaxBlock) k_bluestein(BluesteinArgs<T> b) {
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    constexpr int V = plain_variant<M>();
    constexpr int E = LineShape<M, V>::E;
    constexpr int TT = LineShape<M, V>::T;
    const PencilArgs<T>& a = b.p;
    const int n = a.n;
    const TileInfo ti = decode_tile(a, blockIdx.x, gridDim.x);
    const LineMap lm = line_map(ti.es == 1, a.TW, TT, M);
    const bool valid = lm.l < ti.nvalid;
    const long base = ti.base + (long)lm.l * ti.sb;
    cpx<T>* out = pick(a.arr[0][0], a.arr[0][1], ti.s);
    cpx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int p = lm.t + TT * e;
        cpx<T> x = mk<T>(T(0), T(0));
        if (valid && p < n) x = mul(out[base + (long)p * ti.es], b.chirp[p]);
        v[e] = x;
    }
    fft_canon<T, M, -1, V>(v, lds + lm.off, lm.es, lm.t, a.tw);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = mul(v[e], b.filt[lm.t + TT * e]);
    fft_canon<T, M, +1, V>(v, lds + lm.off, lm.es, lm.t, a.tw);
    if (!valid) return;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int p = lm.t + TT * e;
        if (p < n) out[base + (long)p * ti.es] = mul(v[e], b.chirp[p]);
    }
}

// per-line constants of the update multipliers
struct MultLine {
    int t, TT, nv;
    double invv, fixed2;
};

template <int DEG>
__device__ __forceinline__ double horner(const double* c, double L) {
    double v = c[DEG];
#pragma unroll
    for (int p = DEG - 1; p >= 0; --p) v = v * L + c[p];
    return v;
}

// k-space multipliers of a thread's elements: s1 = m1 - 1 and s2 = m2 / N with
// m1 = A (1 + dt E), m2 = A dt L^src_lap, A = sigma / (1 - dt D).  m1 - 1 is formed as
// (sigma - 1 + dt (sigma E + D)) / (1 - dt D), free of cancellation.  DEG > 0: D and E
// are dense polynomials in L of degree DEG (kernel arguments, or per-sim rows indexed by
// power); DEG = 0: term lists with run-time powers.
template <typename T, int E, int DEG>
__device__ __forceinline__ void update_multipliers(const PencilUpdateArgs<T>& u, const MultLine& ml,
                                                   const double* dsim, const double* lsim, T* s1a, T* s2a,
                                                   double& m1_0, double& m2_0) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int p = ml.t + ml.TT * e;
        const double fv = wrapk(p, ml.nv) * ml.invv;
        const double f2 = ml.fixed2 + fv * fv;
        const double L = -(4.0 * kPi * kPi) * f2;
        double D = 0.0, Ek = 0.0, Lp = 1.0;
        if constexpr (DEG > 0) {
            // (a select of values: a pointer select between global memory and the argument
            // struct would copy the struct to scratch)
            D = dsim ? horner<DEG>(dsim, L) : horner<DEG>(u.dd, L);
            Ek = lsim ? horner<DEG>(lsim, L) : horner<DEG>(u.dl, L);
#pragma unroll
            for (int q = 0; q < DEG; ++q)
                if (q < u.src_lap) Lp *= L;
        } else {
            if (dsim) {
                // per-sim coefficients (global memory), powers shared (kernel argument)
                for (int t = 0; t < u.denom.n; ++t) {
                    double v = dsim[t];
                    for (int q = 0; q < u.denom.pow[t]; ++q) v *= L;
                    D += v;
                }
                for (int t = 0; t < u.lin.n; ++t) {
                    double v = lsim[t];
                    for (int q = 0; q < u.lin.pow[t]; ++q) v *= L;
                    Ek += v;
                }
            } else {
                D = kpoly_eval(u.denom, L);
                Ek = kpoly_eval(u.lin, L);
            }
            for (int q = 0; q < u.src_lap; ++q) Lp *= L;
        }
        const double A0 = recip<T>(1.0 - u.dt * D);
        const double sig = u.filter ? vandeven(u.filter, sqrt(f2) * 2.0 / kPi) : 1.0;
        const double m1m1 = (sig - 1.0 + u.dt * (sig * Ek + D)) * A0;
        const double m2 = sig * A0 * u.dt * Lp;
        if (e == 0) {
            m1_0 = 1.0 + m1m1;
            m2_0 = m2;
        }
        s1a[e] = (T)m1m1;  // the update adds a correction to S: m1 itself is never rounded
        s2a[e] = (T)(m2 * u.scale);
    }
}

// MON: the energy monitor's partial sums are compiled in (a separate instantiation, so
// that the common path does not carry their registers)
// UA is PencilUpdateArgs<T> for the plain kernel (its argument by value: read through a
// reference, the kernel argument costs about ten more VGPRs, one wave per SIMD at 256) and
// const PencilUpdateArgs<T>& for the grouped kernel (arguments read in place from global
// memory: a copy would live in scratch).  EARLY: the line's loads are issued before the
// multipliers are computed (their latency hidden behind that arithmetic, for two more lines
// of registers), for launches whose passes are latency-bound.  MT: 0 the multipliers are
// computed; 1 they are read from u.mtab / u.m0 with the line's loads; 2 they are computed,
// written to u.mtab / u.m0, and the pass ends there (the table's own launch)
template <typename T, int N, bool MON, typename UA = PencilUpdateArgs<T>, bool EARLY = false, int MT = 0>
__device__ __forceinline__ void pencil_update_body(UA u, int blk, int nblk, unsigned char* smem) {
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    constexpr int V = 1;  // low-register schedule: two lines and the multipliers stay in registers
    constexpr int E = LineShape<N, V>::E;
    const PencilArgs<T>& a = u.p;
    const int n = (N > 0) ? N : a.n;
    const int TT = (N > 0) ? LineShape<N, V>::T : n;
    const TileInfo ti = decode_tile(a, blk, nblk);
    const LineMap lm = line_map(ti.es == 1, a.TW, TT, n);
    const bool valid = lm.l < ti.nvalid;
    const long base = ti.base + (long)lm.l * ti.sb;
    cpx<T>* sarr = pick(a.arr[0][0], a.arr[0][1], ti.s);
    cpx<T>* warr = pick(a.arr[1][0], a.arr[1][1], ti.s);
    const double inx = 1.0 / (double)u.g.nx, iny = 1.0 / (double)u.g.ny, inz = 1.0 / (double)u.g.nz;
    // 1) k-space multipliers of this thread's elements, computed before any data is
    //    live to keep register pressure low.  The pencil coordinate p is one of
    //    (kx, ky, kz); the other two are fixed for the line.
    const int bl = ti.b0 + lm.l;
    const int k0 = ksel(ti.ksrc0, ti.kfix0, ti.ia, bl, 0);
    const int k1 = ksel(ti.ksrc1, ti.kfix1, ti.ia, bl, 0);
    const int k2 = ksel(ti.ksrc2, ti.kfix2, ti.ia, bl, 0);
    const int vary = (ti.ksrc0 == 2) ? 0 : (ti.ksrc1 == 2) ? 1 : 2;
    const double f0 = wrapk(k0, u.g.nx) * inx, f1 = wrapk(k1, u.g.ny) * iny;
    const double f2c = (u.g.ndim == 3) ? wrapk(k2, u.g.nz) * inz : 0.0;
    const double fixed2 = (vary == 0 ? 0.0 : f0 * f0) + (vary == 1 ? 0.0 : f1 * f1) + (vary == 2 ? 0.0 : f2c * f2c);
    const int nv = (vary == 0) ? u.g.nx : (vary == 1) ? u.g.ny : u.g.nz;
    const double invv = (vary == 0) ? inx : (vary == 1) ? iny : inz;
    const double wk = (k0 == 0 || 2 * k0 == u.g.nx) ? 1.0 : 2.0;  // Hermitian weight of the kx plane
    // batched sims: this line's sim selects its coefficients and mean.  A line past the end
    // of a tile of the per-sim segment (bl >= nb = number of sims) takes the tile's first
    // sim, so that its discarded multipliers read coefficients inside the tables
    const int sim = (ti.sim_src == 0) ? ti.ia : (ti.sim_src == 1) ? (valid ? bl : ti.b0) : -1;
    const double* dsim = (sim >= 0) ? u.denom_sim + sim * kMaxKTerms : nullptr;
    const double* lsim = (sim >= 0) ? u.lin_sim + sim * kMaxKTerms : nullptr;
    double* mean = (u.mean && sim >= 0) ? u.mean + sim : u.mean;
    // thread holding k = 0 (p = 0 is element 0 of thread 0 of the line)
    const bool has_k0 = valid && mean && lm.t == 0 && (vary == 0 || k0 == 0) && (vary == 1 || k1 == 0) &&
                        (vary == 2 || k2 == 0);
    T s1a[E], s2a[E];
    double m1_0 = 1.0, m2_0 = 0.0;
    cpx<T> gs[E], c[E];
    if constexpr (EARLY) {
        if (u.has_src) load_line<T, E>(gs, warr, base, ti.es, lm.t, TT, valid);
        load_line<T, E>(c, sarr, base, ti.es, lm.t, TT, valid);
    }
    if constexpr (MT == 1) {
        cpx<T> mt[E];
        load_line<T, E>(mt, pick(u.mtab[0], u.mtab[1], ti.s), base, ti.es, lm.t, TT, valid);
        if (has_k0) {
            m1_0 = u.m0[0];
            m2_0 = u.m0[1];
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            s1a[e] = mt[e].x;
            s2a[e] = mt[e].y;
        }
    } else {
        const MultLine ml{lm.t, TT, nv, invv, fixed2};
        // the multipliers and their denominators 1 - dt D(k) are built in FP64 at both
        // precisions; FP32-bulk rounds only the finished s1 = m1 - 1 and s2 to FP32
        if (u.dense == 2) update_multipliers<T, E, 2>(u, ml, dsim, lsim, s1a, s2a, m1_0, m2_0);
        else if (u.dense == 4) update_multipliers<T, E, 4>(u, ml, dsim, lsim, s1a, s2a, m1_0, m2_0);
        else update_multipliers<T, E, 0>(u, ml, dsim, lsim, s1a, s2a, m1_0, m2_0);
    }
    if constexpr (MT == 2) {
        cpx<T> mt[E];
#pragma unroll
        for (int e = 0; e < E; ++e) mt[e] = mk<T>(s1a[e], s2a[e]);
        store_line<T, E>(mt, pick(u.mtab[0], u.mtab[1], ti.s), base, ti.es, lm.t, TT, valid);
        if (has_k0) {
            u.m0[0] = m1_0;
            u.m0[1] = m2_0;
        }
        return;
    }
    // 2) explicit source: forward transform along this axis, scale by its multiplier
    double g0 = 0.0;
    if (u.has_src) {
        if constexpr (!EARLY) load_line<T, E>(gs, warr, base, ti.es, lm.t, TT, valid);
        line_fft<T, N, -1, V>(gs, lds + lm.off, lm.es, n, lm.t, a.tw);
        g0 = (double)gs[0].x;
#pragma unroll
        for (int e = 0; e < E; ++e) gs[e] = mk<T>(s2a[e] * gs[e].x, s2a[e] * gs[e].y);
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) gs[e] = mk<T>(T(0), T(0));
    }
    // 3) semi-implicit update of the state spectrum, which never leaves k-space
    if constexpr (!EARLY) load_line<T, E>(c, sarr, base, ti.es, lm.t, TT, valid);
    double eg = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const cpx<T> cv = c[e];
        if (MON && valid) {
            const double fv = wrapk(lm.t + TT * e, nv) * invv;
            eg += wk * 4.0 * kPi * kPi * (fixed2 + fv * fv) * ((double)cv.x * (double)cv.x + (double)cv.y * (double)cv.y);
        }
        c[e] = mk<T>(cv.x + (s1a[e] * cv.x + gs[e].x), cv.y + (s1a[e] * cv.y + gs[e].y));
    }
    if (has_k0) {
        // k = 0 advances in FP64: the arrays hold c - mean, so their k = 0 coefficient stays 0
        // and mean_new = m1(0) mean + m2(0) G(0) / N
        const double mu = *mean;
        *mean = m1_0 * mu + m2_0 * g0 * u.scale;
        c[0] = mk<T>(T(0), T(0));
    }
    store_line<T, E>(c, sarr, base, ti.es, lm.t, TT, valid);
    if (MON) {
        // the state is FFT(c) / N: |FFT(c)|^2 = |S|^2 / scale^2
        double* red = reinterpret_cast<double*>(lds + lds_elems(n, a.TW));
        const double tot = block_sum(eg, red);
        if (threadIdx.x == 0) u.egrad_partial[blk] = tot / (u.scale * u.scale);
    }
    // 4) the next step's row-pass input: the new state inverse-transformed along this axis
    line_fft<T, N, +1, V>(c, lds + lm.off, lm.es, n, lm.t, a.tw);
    store_line<T, E>(c, warr, base, ti.es, lm.t, TT, valid);
}

template <typename T, int N, bool MON>
__global__ void __launch_bounds__(kMaxBlock) k_pencil_update(PencilUpdateArgs<T> u) {
    extern __shared__ __align__(16) unsigned char smem[];
    pencil_update_body<T, N, MON>(u, blockIdx.x, gridDim.x, smem);
}

// x-rows: t fastest; the half spectrum split needs Z(k) and Z(n-k) from other threads
template <typename T, int E>
__device__ __forceinline__ void split_store(const cpx<T>* v, cpx<T>* L, int n, int t, int TT, long row, bool valid,
                                            cpx<T>* amain, cpx<T>* anyq, cpx<T>* bmain, cpx<T>* bnyq) {
    const int nh = n / 2;
#pragma unroll
    for (int e = 0; e < E; ++e) L[lds_pad(t + TT * e)] = v[e];
    __syncthreads();
    if (valid) {
        for (int k = t; k <= nh; k += TT) {
            const cpx<T> z = L[lds_pad(k)];
            const cpx<T> zm = L[lds_pad((n - k) % n)];
            // A = (z + conj zm) / 2,  B = (z - conj zm) / (2 i)
            const cpx<T> A = mk<T>(T(0.5) * (z.x + zm.x), T(0.5) * (z.y - zm.y));
            const cpx<T> B = mk<T>(T(0.5) * (z.y + zm.y), T(-0.5) * (z.x - zm.x));
            if (k < nh) {
                amain[row * nh + k] = A;
                if (bmain) bmain[row * nh + k] = B;
            } else {
                anyq[row] = A;
                if (bnyq) bnyq[row] = B;
            }
        }
    }
    __syncthreads();
}

// 16-byte access needs 16-byte aligned row bases (every row of an array is, when the array
// base is and nh is even)
__device__ __forceinline__ bool al16(const void* p, const void* q) {
    return ((reinterpret_cast<uintptr_t>(p) | reinterpret_cast<uintptr_t>(q)) & 15) == 0;
}

// split_store for a compiled line shape (N >= 16, E even, so nh = TTc E / 2): the line holds
// Z = A + i B of two real rows; A's half spectrum goes to the row at amain (x-Nyquist value
// at *anyq), B's to bmain / *bnyq when v2.  The k loop is unrolled at compile time, stores use
// immediate offsets from the row bases, and the x-Nyquist pair is written by thread 0.
template <typename T, int E, int TTc>
__device__ __forceinline__ void split_store2(const cpx<T>* v, cpx<T>* L, int n, int t, cpx<T>* amain, cpx<T>* anyq,
                                             cpx<T>* bmain, cpx<T>* bnyq, bool v1, bool v2) {
    const int nh = n / 2;
#pragma unroll
    for (int e = 0; e < E; ++e) L[lds_pad(t + TTc * e)] = v[e];
    __syncthreads();
    constexpr int J = (E * TTc / 2) / TTc;  // = E / 2 k values per thread below nh
    // mirror of k (k <= nh): n - k, and 0 for k = 0
    auto mir = [n](int k) { return k == 0 ? 0 : n - k; };
    bool wide = false;
    if constexpr (sizeof(T) == 4 && J % 2 == 0) wide = al16(amain, v2 ? bmain : amain);
    if constexpr (sizeof(T) == 4 && J % 2 == 0) {
        if (wide) {
            // FP32: lane t stores k = 2t, 2t + 1 (+ 2 TTc j) of each row as one 16-byte value
            float4* fa = reinterpret_cast<float4*>(amain) + t;
            float4* fb = reinterpret_cast<float4*>(bmain) + t;
            float4 A4[J / 2], B4[J / 2];
#pragma unroll
            for (int j = 0; j < J / 2; ++j) {
                const int k = 2 * t + 2 * TTc * j;
                const cpx<T> z0 = L[lds_pad(k)], z1 = L[lds_pad(k + 1)];
                const cpx<T> m0 = L[lds_pad(mir(k))], m1 = L[lds_pad(n - k - 1)];
                A4[j] = make_float4(0.5f * (z0.x + m0.x), 0.5f * (z0.y - m0.y), 0.5f * (z1.x + m1.x),
                                    0.5f * (z1.y - m1.y));
                B4[j] = make_float4(0.5f * (z0.y + m0.y), -0.5f * (z0.x - m0.x), 0.5f * (z1.y + m1.y),
                                    -0.5f * (z1.x - m1.x));
            }
            if (v1 && v2) {
#pragma unroll
                for (int j = 0; j < J / 2; ++j) {
                    fa[TTc * j] = A4[j];
                    fb[TTc * j] = B4[j];
                }
            } else if (v1) {
#pragma unroll
                for (int j = 0; j < J / 2; ++j) fa[TTc * j] = A4[j];
            }
        }
    }
    if (!wide) {
        cpx<T>* a1 = amain + t;
        cpx<T>* b1 = bmain + t;
        cpx<T> A[J], B[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int k = t + TTc * j;
            const cpx<T> z = L[lds_pad(k)];
            const cpx<T> zm = L[lds_pad(mir(k))];
            // A = (z + conj zm) / 2,  B = (z - conj zm) / (2 i)
            A[j] = mk<T>(T(0.5) * (z.x + zm.x), T(0.5) * (z.y - zm.y));
            B[j] = mk<T>(T(0.5) * (z.y + zm.y), T(-0.5) * (z.x - zm.x));
        }
        if (v1 && v2) {
            // every line but possibly the last: one branch for all stores
#pragma unroll
            for (int j = 0; j < J; ++j) {
                a1[TTc * j] = A[j];
                b1[TTc * j] = B[j];
            }
        } else if (v1) {
#pragma unroll
            for (int j = 0; j < J; ++j) a1[TTc * j] = A[j];
        }
    }
    if (t == 0) {
        const cpx<T> z = L[lds_pad(nh)];
        if (v1) *anyq = mk<T>(z.x, T(0));
        if (v2) *bnyq = mk<T>(z.y, T(0));
    }
    __syncthreads();
}

// split_store2 for the fused row pass: rows r1 (A) and r1 + 1 (B) of one array
template <typename T, int E, int TTc>
__device__ __forceinline__ void split_store_rows(const cpx<T>* v, cpx<T>* L, int n, int t, long r1, bool v1, bool v2,
                                                 cpx<T>* gmain, cpx<T>* gnyq) {
    const int nh = n / 2;
    split_store2<T, E, TTc>(v, L, n, t, gmain + r1 * nh, gnyq + r1, gmain + (r1 + 1) * nh, gnyq + r1 + 1, v1, v2);
}

// Hermitian extension of a half-spectrum row at canonical positions
template <typename T, int E>
__device__ __forceinline__ void load_herm(cpx<T>* v, const cpx<T>* __restrict__ m, const cpx<T>* __restrict__ q,
                                          long row, int n, int t, int TT, bool valid) {
    const int nh = n / 2;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int p = t + TT * e;
        cpx<T> x = mk<T>(T(0), T(0));
        if (valid) {
            if (p < nh) x = m[row * nh + p];
            else if (p == nh) x = q[row];
            else if (n - p == nh) x = conjg(q[row]);  // odd n: the mirror is the stored column kx = nh
            else x = conjg(m[row * nh + (n - p)]);
        }
        v[e] = x;
    }
}

// rows r1 and r1 + 1 (when valid) of a half-spectrum array, Hermitian-extended and
// packed as A + i B, so that one inverse transform yields both real rows.  Element e of
// thread t is position p = t + TT e: below nh it is read directly, above nh it is the
// conjugate of position n - p.  With a compiled length (N > 0, TT = TTc) the region of
// each e is known at compile time, so the loads use per-row base pointers with immediate
// offsets; only the element that straddles nh tests at run time.
template <typename T, int E, int TTc>
__device__ __forceinline__ void load_herm_pair(cpx<T>* v, const cpx<T>* m, const cpx<T>* q, long r1, int n, int t,
                                               int TT, bool v1, bool v2) {
    const int nh = n / 2;
    const cpx<T>* a1 = m + r1 * nh + t;        // direct: row r1 at position t + TT e
    const cpx<T>* b1 = m + r1 * nh + (n - t);  // mirror: row r1 at position n - t - TT e
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int lo = (TTc > 0 ? TTc : TT) * e;
        const int hi = lo + (TTc > 0 ? TTc : TT) - 1;
        const int p = t + lo;
        cpx<T> A = mk<T>(T(0), T(0)), B = A;
        if (TTc > 0 && hi < nh) {
            if (v1) A = a1[lo];
            if (v2) B = a1[lo + nh];
        } else if (TTc > 0 && lo > nh) {
            if (v1) A = conjg(b1[-lo]);
            if (v2) B = conjg(b1[nh - lo]);
        } else if (p < nh) {
            if (v1) A = a1[lo];
            if (v2) B = a1[lo + nh];
        } else if (p == nh) {
            if (v1) A = q[r1];
            if (v2) B = q[r1 + 1];
        } else if (n - p == nh) {
            // odd n: the mirror of p is the stored column kx = nh, a frequency of its own
            if (v1) A = conjg(q[r1]);
            if (v2) B = conjg(q[r1 + 1]);
        } else {
            if (v1) A = conjg(b1[-lo]);
            if (v2) B = conjg(b1[nh - lo]);
        }
        v[e] = mk<T>(A.x - B.y, A.y + B.x);
    }
}

// Two half-spectrum rows (A: main row at pa, x-Nyquist value at *qa; B: pb, *qb),
// Hermitian-extended and packed as A + i B, so that one inverse transform yields both real
// rows.  Each half-spectrum value is read from HBM once: the direct positions are loaded and
// staged in the line's LDS (row A at [0, nh], row B at [nh + 1, n + 1]), and the mirrored
// positions are read back from LDS.  The line's LDS region (lds_line(n) >= n + 2 elements
// for n >= 16) is free again on return.  Every lane issues all of its loads before it uses
// any of them; the row pointers must be valid addresses even when v1 / v2 is false (the
// values are then discarded), so no load sits behind a branch.
template <typename T, int E, int TTc>
__device__ __forceinline__ void load_herm2_lds(cpx<T>* v, const cpx<T>* pa, const cpx<T>* qa, const cpx<T>* pb,
                                               const cpx<T>* qb, int n, int t, bool v1, bool v2, cpx<T>* L) {
    const int nh = n / 2;
    if constexpr (sizeof(T) == 4 && (TTc * E) % (4 * TTc) == 0) {
        if (al16(pa, pb)) {
            // FP32: lane t loads positions 2t, 2t + 1 (+ 2 TTc e) of each row as one 16-byte
            // value; every canonical element then comes from LDS
            constexpr int NL = (TTc * E / 2) / (2 * TTc);  // 16-byte loads per row below nh
            const float4* fa = reinterpret_cast<const float4*>(pa) + t;
            const float4* fb = reinterpret_cast<const float4*>(pb) + t;
            float4 A4[NL], B4[NL];
#pragma unroll
            for (int e = 0; e < NL; ++e) {
                A4[e] = fa[TTc * e];
                B4[e] = fb[TTc * e];
            }
            const cpx<T> qav = *qa, qbv = *qb;
            const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int e = 0; e < NL; ++e) {
                const int p = 2 * t + 2 * TTc * e;
                const float4 a4 = v1 ? A4[e] : z4, b4 = v2 ? B4[e] : z4;
                L[p] = mk<T>(a4.x, a4.y);
                L[p + 1] = mk<T>(a4.z, a4.w);
                L[nh + 1 + p] = mk<T>(b4.x, b4.y);
                L[nh + 2 + p] = mk<T>(b4.z, b4.w);
            }
            if (t == 0) {
                L[nh] = v1 ? qav : mk<T>(T(0), T(0));
                L[n + 1] = v2 ? qbv : mk<T>(T(0), T(0));
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int p = t + TTc * e;
                cpx<T> a, b;
                if (p <= nh) {
                    a = L[p];
                    b = L[nh + 1 + p];
                } else {
                    a = conjg(L[n - p]);
                    b = conjg(L[nh + 1 + n - p]);
                }
                v[e] = mk<T>(a.x - b.y, a.y + b.x);
            }
            __syncthreads();
            return;
        }
    }
    const cpx<T> zero = mk<T>(T(0), T(0));
    cpx<T> A[E], B[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int lo = TTc * e, hi = lo + TTc - 1;
        if (lo > nh) continue;  // mirrored for every lane
        const int p = t + lo;
        if (hi < nh) {
            A[e] = pa[p];
            B[e] = pb[p];
        } else {
            // the element that straddles nh: direct below it, the x-Nyquist value at it
            const cpx<T>* sa = (p < nh) ? pa + p : (p == nh ? qa : pa);
            const cpx<T>* sb = (p < nh) ? pb + p : (p == nh ? qb : pb);
            A[e] = *sa;
            B[e] = *sb;
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int lo = TTc * e, hi = lo + TTc - 1;
        if (lo > nh) continue;
        const int p = t + lo;
        const cpx<T> a = v1 ? A[e] : zero, b = v2 ? B[e] : zero;
        if (hi < nh || p <= nh) {
            L[p] = a;
            L[nh + 1 + p] = b;
            v[e] = mk<T>(a.x - b.y, a.y + b.x);
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int lo = TTc * e, hi = lo + TTc - 1;
        const int p = t + lo;
        if (hi < nh) continue;
        if (lo > nh || p > nh) {
            const cpx<T> a = conjg(L[n - p]), b = conjg(L[nh + 1 + n - p]);
            v[e] = mk<T>(a.x - b.y, a.y + b.x);
        }
    }
    __syncthreads();
}

// load_herm2_lds for the fused row pass: rows r1 (A) and r1 + 1 (B) of one array; a row
// past the end reads a valid row and is discarded
template <typename T, int E, int TTc>
__device__ __forceinline__ void load_herm_pair_lds(cpx<T>* v, const cpx<T>* m, const cpx<T>* q, long r1, int n, int t,
                                                   bool v1, bool v2, cpx<T>* L) {
    const int nh = n / 2;
    const long ra = v1 ? r1 : 0, rb = v2 ? r1 + 1 : ra;
    load_herm2_lds<T, E, TTc>(v, m + ra * nh, q + ra, m + rb * nh, q + rb, n, t, v1, v2, L);
}

// sum_ij dc[i][j] x0^i x1^j in T, Horner in x0 of polynomials in x1 (the k_derived formula)
template <typename T>
__device__ __forceinline__ T dense_poly(const double (&dc)[4][4], T x0, T x1) {
    T acc = T(0);
#pragma unroll
    for (int p = 3; p >= 0; --p) {
        const T inner = (((T)dc[p][3] * x1 + (T)dc[p][2]) * x1 + (T)dc[p][1]) * x1 + (T)dc[p][0];
        acc = acc * x0 + inner;
    }
    return acc;
}

// the dense derived-field polynomial of array q at element off
template <typename T>
__device__ __forceinline__ T row_poly(const RowR2CArgs<T>& a, int q, long off) {
    const T x0 = a.pf0[q][off];
    const T x1 = a.pf1[q] ? a.pf1[q][off] : T(0);
    return dense_poly<T>(a.dc[q], x0, x1);
}

// Row kernels of the general path.  Compiled lengths n >= 16 (kStaged) stage each line's
// Hermitian loads and split stores through its LDS region (every half-spectrum value read
// and written once, 16-byte FP32 accesses) and keep the block's copy of the twiddle table in
// LDS after the exchange lines; rows past the end read row 0 and store nothing.
template <int N, int E>
constexpr bool row_staged() {
    return N >= 16 && E % 2 == 0;
}

// the block's copy of the n-point twiddle table, after the exchange lines of TW lines
template <typename T>
__device__ __forceinline__ const cpx<T>* stage_twiddles(cpx<T>* lds, const cpx<T>* tw, int n, int TW) {
    cpx<T>* twl = lds + lds_elems(n, TW);
    for (int i = threadIdx.x; i < n; i += blockDim.x) twl[i] = tw[i];
    __syncthreads();
    return twl;
}

template <typename T, int N>
__global__ void __launch_bounds__(kMaxBlock) k_rows_r2c(RowR2CArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    constexpr int V = plain_variant<N>();
    constexpr int E = LineShape<N, V>::E;
    constexpr bool kStaged = row_staged<N, E>();
    const int n = (N > 0) ? N : a.nx;
    const int nh = n / 2;
    const int TT = (N > 0) ? LineShape<N, V>::T : n;
    const LineMap lm = line_map(true, a.TW, TT, n);
    const long row = (long)blockIdx.x * a.TW + lm.l;
    const bool valid = row < a.rows;
    const long rv = valid ? row : 0;
    const cpx<T>* twp = a.tw;
    if constexpr (kStaged) twp = stage_twiddles<T>(lds, a.tw, n, a.TW);
#pragma unroll
    for (int q = 0; q < kMaxArr; q += 2) {
        if (q >= a.narr) break;
        const T* ia = a.in[q];
        const T* ib = (q + 1 < a.narr) ? a.in[q + 1] : nullptr;
        const bool two = q + 1 < a.narr;
        cpx<T> v[E];
        // the source rows of the pair: (f0, f1) of a polynomial or the plain input; when the
        // second array reads the same rows as the first (two polynomials of the same fields),
        // each row is loaded once, and every load is issued before any arithmetic
        const T* s0 = a.poly[q] ? a.pf0[q] : ia;
        const T* s1 = a.poly[q] ? a.pf1[q] : nullptr;
        const T* t0 = two ? (a.poly[q + 1] ? a.pf0[q + 1] : ib) : nullptr;
        const T* t1 = (two && a.poly[q + 1]) ? a.pf1[q + 1] : nullptr;
        if (!two || (t0 == s0 && t1 == s1)) {
            T x0[E], x1[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const long off = rv * n + lm.t + TT * e;
                x0[e] = s0 ? s0[off] : T(0);
                x1[e] = s1 ? s1[off] : T(0);
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const T x = a.poly[q] ? dense_poly<T>(a.dc[q], x0[e], x1[e]) : x0[e];
                const T y = !two ? T(0) : a.poly[q + 1] ? dense_poly<T>(a.dc[q + 1], x0[e], x1[e]) : x0[e];
                v[e] = valid ? mk<T>(x * a.in_scale, y * a.in_scale) : mk<T>(T(0), T(0));
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const long off = rv * n + lm.t + TT * e;
                T x = T(0), y = T(0);
                if (a.poly[q]) x = row_poly(a, q, off) * a.in_scale;
                else if (ia) x = ia[off] * a.in_scale;
                if (two && a.poly[q + 1]) y = row_poly(a, q + 1, off) * a.in_scale;
                else if (ib) y = ib[off] * a.in_scale;
                v[e] = valid ? mk<T>(x, y) : mk<T>(T(0), T(0));
            }
        }
        line_fft<T, N, -1, V>(v, lds + lm.off, 1, n, lm.t, twp);
        if constexpr (kStaged) {
            constexpr int TTc = LineShape<N, V>::T;
            cpx<T>* am = a.main[q] + rv * nh;
            split_store2<T, E, TTc>(v, lds + lm.off, n, lm.t, am, a.nyq[q] + rv, two ? a.main[q + 1] + rv * nh : am,
                                    two ? a.nyq[q + 1] + rv : a.nyq[q] + rv, valid, valid && two);
        } else {
            split_store<T, E>(v, lds + lm.off, n, lm.t, TT, row, valid, a.main[q], a.nyq[q],
                              two ? a.main[q + 1] : nullptr, two ? a.nyq[q + 1] : nullptr);
        }
    }
}

template <typename T, int N>
__global__ void __launch_bounds__(kMaxBlock) k_rows_c2r(RowC2RArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    constexpr int V = plain_variant<N>();
    constexpr int E = LineShape<N, V>::E;
    constexpr bool kStaged = row_staged<N, E>();
    const int n = (N > 0) ? N : a.nx;
    const int nh = n / 2;
    const int TT = (N > 0) ? LineShape<N, V>::T : n;
    const LineMap lm = line_map(true, a.TW, TT, n);
    const long row = (long)blockIdx.x * a.TW + lm.l;
    const bool valid = row < a.rows;
    const long rv = valid ? row : 0;
    const double add0 = a.add_mean ? *a.add_mean : 0.0;
    const cpx<T>* twp = a.tw;
    if constexpr (kStaged) twp = stage_twiddles<T>(lds, a.tw, n, a.TW);
#pragma unroll
    for (int q = 0; q < kMaxArr; q += 2) {
        if (q >= a.narr) break;
        const bool two = q + 1 < a.narr;
        cpx<T> v[E];
        if constexpr (kStaged) {
            constexpr int TTc = LineShape<N, V>::T;
            const cpx<T>* pa = a.main[q] + rv * nh;
            load_herm2_lds<T, E, TTc>(v, pa, a.nyq[q] + rv, two ? a.main[q + 1] + rv * nh : pa,
                                      two ? a.nyq[q + 1] + rv : a.nyq[q] + rv, n, lm.t, valid, valid && two,
                                      lds + lm.off);
        } else {
            cpx<T> A[E];
            load_herm<T, E>(A, a.main[q], a.nyq[q], row, n, lm.t, TT, valid);
            if (two) {
                load_herm<T, E>(v, a.main[q + 1], a.nyq[q + 1], row, n, lm.t, TT, valid);
#pragma unroll
                for (int e = 0; e < E; ++e) v[e] = mk<T>(A[e].x - v[e].y, A[e].y + v[e].x);  // A + i B
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e) v[e] = A[e];
            }
        }
        line_fft<T, N, +1, V>(v, lds + lm.off, 1, n, lm.t, twp);
        if (valid) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const long off = row * n + lm.t + TT * e;
                const T r = v[e].x * a.scale;
                a.out[q][off] = (q == 0 && a.add_mean) ? (T)((double)r + add0) : r;
                if (two) a.out[q + 1][off] = v[e].y * a.scale;
            }
        }
    }
}

// MODE: 0 = at most two derived arrays (one transform pair, the derived rows formed in place
// of the real rows), 1 = the elastic source, 2 = the adjoint's lambda, 3 = more than two
// derived arrays; separate instantiations keep the common mode 0 at its own register count
template <typename T, int N, int MODE>
__global__ void __launch_bounds__(kMaxBlock) k_rows_sync(RowSyncArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    constexpr int V = plain_variant<N>();
    constexpr int E = LineShape<N, V>::E;
    constexpr bool kStaged = row_staged<N, E>();
    constexpr int TTc = (N > 0) ? LineShape<N, V>::T : 0;
    const int n = (N > 0) ? N : a.nx;
    const int nh = n / 2;
    const int TT = (N > 0) ? LineShape<N, V>::T : n;
    const LineMap lm = line_map(true, a.TW, TT, n);
    const long row = (long)blockIdx.x * a.TW + lm.l;
    const bool valid = row < a.rows;
    const long rv = valid ? row : 0;
    const bool two = a.fmain[1] != nullptr;
    const cpx<T>* twp = a.tw;
    if constexpr (kStaged) twp = stage_twiddles<T>(lds, a.tw, n, a.TW);
    // one derived row (A only) or a pair (A, B) of the forward transform in v to its arrays
    auto store_pair = [&](const cpx<T>* vv, cpx<T>* am, cpx<T>* aq, cpx<T>* bm, cpx<T>* bq) {
        if constexpr (kStaged) {
            cpx<T>* a0 = am + rv * nh;
            split_store2<T, E, TTc>(vv, lds + lm.off, n, lm.t, a0, aq + rv, bm ? bm + rv * nh : a0,
                                    bq ? bq + rv : aq + rv, valid, valid && bm != nullptr);
        } else {
            split_store<T, E>(vv, lds + lm.off, n, lm.t, TT, row, valid, am, aq, bm, bq);
        }
    };
    cpx<T> v[E];
    if constexpr (kStaged) {
        const cpx<T>* pa = a.fmain[0] + rv * nh;
        load_herm2_lds<T, E, TTc>(v, pa, a.fnyq[0] + rv, two ? a.fmain[1] + rv * nh : pa,
                                  two ? a.fnyq[1] + rv : a.fnyq[0] + rv, n, lm.t, valid, valid && two, lds + lm.off);
    } else {
        cpx<T> A[E];
        load_herm<T, E>(A, a.fmain[0], a.fnyq[0], row, n, lm.t, TT, valid);
        if (two) {
            load_herm<T, E>(v, a.fmain[1], a.fnyq[1], row, n, lm.t, TT, valid);
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = mk<T>(A[e].x - v[e].y, A[e].y + v[e].x);  // A + i B
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = A[e];
        }
    }
    line_fft<T, N, +1, V>(v, lds + lm.off, 1, n, lm.t, twp);
    if constexpr (MODE == 0) {
        // the real rows (rounded to T as the inverse pass stores them) and, in place, the
        // derived rows of the (one or two) dense polynomials
        const bool dtwo = a.nd > 1;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const T x0 = v[e].x * a.scale;
            const T x1 = two ? v[e].y * a.scale : T(0);
            if (valid) {
                const long off = row * n + lm.t + TT * e;
                if (a.fout[0]) a.fout[0][off] = x0;
                if (two && a.fout[1]) a.fout[1][off] = x1;
            }
            const T d0 = valid ? dense_poly<T>(a.dc[0], x0, x1) : T(0);
            const T d1 = (valid && dtwo) ? dense_poly<T>(a.dc[1], x0, x1) : T(0);
            v[e] = mk<T>(d0, d1);
        }
        line_fft<T, N, -1, V>(v, lds + lm.off, 1, n, lm.t, twp);
        store_pair(v, a.dmain[0], a.dnyq[0], dtwo ? a.dmain[1] : nullptr, dtwo ? a.dnyq[1] : nullptr);
        return;
    }
    // the real rows, rounded to T as the inverse pass stores them
    T x0[E], x1[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        x0[e] = v[e].x * a.scale;
        x1[e] = two ? v[e].y * a.scale : T(0);
        if (valid && MODE == 3) {
            const long off = row * n + lm.t + TT * e;
            if (a.fout[0]) a.fout[0][off] = x0[e];
            if (two && a.fout[1]) a.fout[1][off] = x1[e];
        }
    }
    if constexpr (MODE == 2) {
        // the k_adj_lambda formula on the inverse rows
#pragma unroll
        for (int e = 0; e < E; ++e) {
            T lam = T(0);
            if (valid) {
                const long off = row * n + lm.t + TT * e;
                double d = 0.0;
                if (a.phi) {
                    const double x = (double)a.phi[off];
                    for (int k = a.npoly - 1; k >= 0; --k) d = d * x + a.gd[k];
                }
                lam = (T)((double)x0[e] + d * (double)x1[e]);
                a.fout[0][off] = lam;
            }
            v[e] = mk<T>(lam, T(0));
        }
        line_fft<T, N, -1, V>(v, lds + lm.off, 1, n, lm.t, twp);
        store_pair(v, a.dmain[0], a.dnyq[0], nullptr, nullptr);
        return;
    }
    if constexpr (MODE == 1) {
        // P = H'(phi) (S - 2 e* H(phi)) in T on the inverse row, H = 3 phi^2 - 2 phi^3
        // (homoLinElast.go:10-12, 91-127)
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const T x = a.phi[rv * n + lm.t + TT * e];
            const T H = T(3) * x * x - T(2) * x * x * x;
            const T dH = T(6) * x - T(6) * x * x;
            const T P = valid ? dH * (x0[e] - (T)a.two_estar * H) : T(0);
            v[e] = mk<T>(P, T(0));
        }
        line_fft<T, N, -1, V>(v, lds + lm.off, 1, n, lm.t, twp);
        store_pair(v, a.dmain[0], a.dnyq[0], nullptr, nullptr);
        return;
    }
    if constexpr (MODE == 3) {
#pragma unroll
        for (int q = 0; q < kMaxArr; q += 2) {
            if (q >= a.nd) break;
            const bool dtwo = q + 1 < a.nd;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const T d0 = valid ? dense_poly<T>(a.dc[q], x0[e], x1[e]) : T(0);
                const T d1 = (valid && dtwo) ? dense_poly<T>(a.dc[q + 1], x0[e], x1[e]) : T(0);
                v[e] = mk<T>(d0, d1);
            }
            line_fft<T, N, -1, V>(v, lds + lm.off, 1, n, lm.t, twp);
            store_pair(v, a.dmain[q], a.dnyq[q], dtwo ? a.dmain[q + 1] : nullptr, dtwo ? a.dnyq[q + 1] : nullptr);
        }
    }
}

// real space of the row pass (row r1 in x, row r1 + 1 in y): g from the shifted
// coefficients with DEG of them (those beyond npoly are zero), optional real-space
// output c_out and the monitor's FP64 partial sums of the energy density (nbulk
// coefficients fb, held in registers: read from the arguments inside this loop, they would
// be read again after every store where the arguments are in global memory)
template <typename T, int E, int DEG, bool MON>
__device__ __forceinline__ void rows_pointwise(cpx<T>* v, const T* gp, T* c_out, const double* fb, int nbulk, long r1,
                                               int n, int t, int TT, bool v1, bool v2, double mu, double& sb,
                                               double& sm) {
    const T muT = (T)mu;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const T d1 = v[e].x, d2 = v[e].y;
        T g1 = gp[DEG - 1], g2 = gp[DEG - 1];
#pragma unroll
        for (int q = DEG - 2; q >= 0; --q) {
            g1 = g1 * d1 + gp[q];
            g2 = g2 * d2 + gp[q];
        }
        v[e] = mk<T>(g1, g2);
        const long off = r1 * n + t + TT * e;
        if (c_out) {
            if (v1) c_out[off] = d1 + muT;
            if (v2) c_out[off + n] = d2 + muT;
        }
        if (MON) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (h == 0 ? !v1 : !v2) continue;
                const double cd = (double)(h == 0 ? d1 : d2) + mu;
                // Horner over all kMaxPoly coefficients (zero from nbulk on): the leading
                // zero terms leave f = 0 * cd + 0 until the first nonzero one, exactly as a
                // start at q = nbulk - 1 would, with no branch per coefficient
                double f = 0.0;
#pragma unroll
                for (int q = kMaxPoly - 1; q >= 0; --q) f = f * cd + fb[q];
                sb += f;
                sm += cd;
            }
        }
    }
}

// V: transform schedule (0 plain, 1 low-register); MON: monitor partial sums compiled in.
// RA is RowFusedArgs<T> (by value) for the plain kernel and const RowFusedArgs<T>& where the
// arguments are read in place from global memory (the persistent step)
template <typename T, int N, int V, bool MON, typename RA = RowFusedArgs<T>>
__device__ __forceinline__ void rows_fused_body(RA a, int blk, unsigned char* smem) {
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    constexpr int E = LineShape<N, V>::E;
    const int n = (N > 0) ? N : a.nx;
    const int TT = (N > 0) ? LineShape<N, V>::T : n;
    const LineMap lm = line_map(true, a.TW, TT, n);
    // this line carries rows r1 and r1 + 1
    const long r1 = 2 * ((long)blk * a.TW + lm.l);
    const bool v1 = r1 < a.rows, v2 = r1 + 1 < a.rows;
    const long sim = (a.rows_per_sim > 0 && v1) ? r1 / a.rows_per_sim : -1;
    const double mu = a.mean ? a.mean[sim >= 0 ? sim : 0] : 0.0;
    // g about the mean: g(mu + dc) = sum_q gp[q] dc^q, the Taylor shift of the coefficients
    // formed in FP64 (repeated synthetic division), so that the bulk evaluation rounds
    // relative to the fluctuation dc rather than to c
    T gp[kMaxPoly];
    {
        double gd[kMaxPoly];
#pragma unroll
        for (int q = 0; q < kMaxPoly; ++q)
            gd[q] = (q < a.npoly) ? ((sim >= 0) ? a.gpoly_sim[sim * kMaxPoly + q] : a.gpoly[q]) : 0.0;
        if (a.mean) {
#pragma unroll
            for (int i = 0; i < kMaxPoly - 1; ++i)
#pragma unroll
                for (int j = kMaxPoly - 2; j >= i; --j) gd[j] += mu * gd[j + 1];
        }
#pragma unroll
        for (int q = 0; q < kMaxPoly; ++q) gp[q] = (T)gd[q];
    }
    // the block's copy of the twiddle table: the transforms' per-stage twiddles are then
    // LDS reads rather than gathers from global memory
    cpx<T>* twl = lds + lds_elems(n, a.TW);
    for (int i = threadIdx.x; i < n; i += blockDim.x) twl[i] = a.tw[i];
    __syncthreads();
    cpx<T> v[E];
    if constexpr (N >= 16)
        load_herm_pair_lds<T, E, LineShape<N, V>::T>(v, a.s_main, a.s_nyq, r1, n, lm.t, v1, v2, lds + lm.off);
    else
        load_herm_pair<T, E, (N > 0 ? LineShape<N, V>::T : 0)>(v, a.s_main, a.s_nyq, r1, n, lm.t, TT, v1, v2);
    line_fft<T, N, +1, V>(v, lds + lm.off, 1, n, lm.t, twl);
    // per-thread partials of at most 2 E values: plain FP64 sums (the blocks' partials are
    // added with compensation)
    double sb = 0.0, sm = 0.0;
    double fb[kMaxPoly];
    const int nbulk = MON ? a.nbulk : 0;
#pragma unroll
    for (int q = 0; q < kMaxPoly; ++q) fb[q] = (MON && q < nbulk) ? a.bulk[q] : 0.0;
    T* const c_out = a.c_out;
    if (a.npoly <= 4) rows_pointwise<T, E, 4, MON>(v, gp, c_out, fb, nbulk, r1, n, lm.t, TT, v1, v2, mu, sb, sm);
    else rows_pointwise<T, E, kMaxPoly, MON>(v, gp, c_out, fb, nbulk, r1, n, lm.t, TT, v1, v2, mu, sb, sm);
    if (MON) {
        double* red = reinterpret_cast<double*>(lds + lds_elems(n, a.TW) + n);
        const double b = block_sum(sb, red);
        const double ms = block_sum(sm, red);
        if (threadIdx.x == 0) {
            a.bulk_partial[blk] = b;
            a.mass_partial[blk] = ms;
        }
        if (a.series && blk == 0)
            series_slot(a.prev_bulk, a.prev_mass, a.prev_nb, a.prev_eg, a.prev_ne, a.prev_ceg, a.series, a.counter);
    }
    line_fft<T, N, -1, V>(v, lds + lm.off, 1, n, lm.t, twl);
    if constexpr (row_staged<N, E>())
        split_store_rows<T, E, LineShape<N, V>::T>(v, lds + lm.off, n, lm.t, r1, v1, v2, a.g_main, a.g_nyq);
    else
        split_store<T, E>(v, lds + lm.off, n, lm.t, TT, r1, v1, a.g_main, a.g_nyq, v2 ? a.g_main + n / 2 : nullptr,
                          v2 ? a.g_nyq + 1 : nullptr);
}

template <typename T, int N, int V, bool MON>
__global__ void __launch_bounds__(kMaxBlock) k_rows_fused(RowFusedArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    rows_fused_body<T, N, V, MON>(a, blockIdx.x, smem);
}

// ---------------------------------------------------------------------------
// grouped launches of a batch: one launch advances every size class.  The classes' block
// ranges follow one another, each starting at a multiple of 8 so that the XCD-aware tile
// order of a class matches the XCD of its blocks (block b runs on XCD b mod 8); blocks past
// a class's own count (the padding) exit at once.  A class's arguments come from a device
// array, read uniformly by the block; its compiled length is selected by a switch.
// ---------------------------------------------------------------------------
#define GPF_GROUP_LENGTHS(X) X(64) X(128) X(256) X(512) X(1024)

// class of block b and b's index inside the class (unrolled selects: the argument struct is
// never indexed dynamically)
template <typename T>
__device__ __forceinline__ int group_class(const BatchGroupArgs<T>& g, int b, int& local, int& count) {
    int c = 0, first = 0, cnt = g.blocks[0];
#pragma unroll
    for (int k = 1; k < kMaxGroup; ++k)
        if (k < g.ncls && b >= g.first[k]) {
            c = k;
            first = g.first[k];
            cnt = g.blocks[k];
        }
    local = b - first;
    count = cnt;
    return c;
}

template <typename T>
__global__ void __launch_bounds__(kMaxBlock) k_rows_fused_group(BatchGroupArgs<T> g) {
    extern __shared__ __align__(16) unsigned char smem[];
    int blk, cnt;
    const int c = group_class(g, blockIdx.x, blk, cnt);
    if (blk >= cnt) return;
    const RowFusedArgs<T> a = g.rows[c];
    switch (a.nx * 2 + a.variant) {
#define GPF_G(NN)                                                                           \
    case NN * 2 + 1: rows_fused_body<T, NN, 1, false>(a, blk, smem); break;                 \
    case NN * 2:                                                                           \
        if constexpr (plain_variant<NN>() == 0) rows_fused_body<T, NN, 0, false>(a, blk, smem); \
        break;
        GPF_GROUP_LENGTHS(GPF_G)
#undef GPF_G
        default: break;
    }
}

template <typename T>
__global__ void __launch_bounds__(kMaxBlock) k_pencil_update_group(BatchGroupArgs<T> g) {
    extern __shared__ __align__(16) unsigned char smem[];
    int blk, cnt;
    const int c = group_class(g, blockIdx.x, blk, cnt);
    if (blk >= cnt) return;
    // read in place (a copy of the argument struct would live in scratch memory)
    const PencilUpdateArgs<T>& u = g.upd[c];
    switch (u.p.n) {
#define GPF_G(NN) \
    case NN: pencil_update_body<T, NN, false, const PencilUpdateArgs<T>&>(u, blk, cnt, smem); break;
        GPF_GROUP_LENGTHS(GPF_G)
#undef GPF_G
        default: break;
    }
}

// v: 0 = plain pass, 1 = update pass
int threads_per_line(int n, int v = 0) {
    switch (n) {
#define GPF_T(NN) case NN: return (v || plain_variant<NN>()) ? Sched<NN, 1>::T : Sched<NN>::T;
        GPF_FAST_LENGTHS(GPF_T)
#undef GPF_T
        default: return n;  // brute-force DFT: one thread per output
    }
}

// Length sets of the launch functions: 0 = the core lengths and the direct DFT (n <= 64),
// 1 = the extended lengths.  Each set's kernels are compiled in units of their own: a code
// object holding every length made each dispatch of the 256^3 step about 1 us slower (2-4 %
// of the step), with the kernels themselves unchanged.
__host__ __device__ constexpr bool ext_length(int n) {
    switch (n) {
#define GPF_C(NN) case NN:
        GPF_EXT_LENGTHS(GPF_C)
#undef GPF_C
        return true;
        default:
            return false;
    }
}

template <int SET, typename F>
hipError_t dispatch_size(int n, F&& f) {
    if constexpr (SET == 0) {
        switch (n) {
#define GPF_D(NN) case NN: return f(std::integral_constant<int, NN>{});
            GPF_CORE_LENGTHS(GPF_D)
#undef GPF_D
            default:
                if (n >= 1 && n <= 64) return f(std::integral_constant<int, 0>{});
                return hipErrorInvalidValue;
        }
    } else {
        switch (n) {
#define GPF_D(NN) case NN: return f(std::integral_constant<int, NN>{});
            GPF_EXT_LENGTHS(GPF_D)
#undef GPF_D
            default:
                return hipErrorInvalidValue;
        }
    }
}

template <typename K, typename A>
hipError_t launch_k(K kernel, dim3 grid, dim3 block, size_t smem, hipStream_t s, const A& args) {
    if (grid.x == 0) return hipSuccess;
    if (smem > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kernel, grid, block, smem, s, args);
    return hipGetLastError();
}

}  // namespace

// the pencils of one direction (the two directions are instantiated in separate units), for
// the lengths of set SET
template <typename T, int DIR, int SET>
hipError_t launch_pencils_dir_set(const PencilArgs<T>& a0, int n, hipStream_t s) {
    constexpr int dir = DIR;
    PencilArgs<T> a = a0;
    a.n = n;
    if (a.lpt != 1 && a.lpt != 2) return hipErrorInvalidValue;
    if (sizeof(T) == 8) a.lpt = 1;  // FP64 lines are 16-byte values already; two per thread would spill
    const long blocks = pencil_blocks(a.seg, a.nseg, a.TW * a.lpt, &a.tiles0);
    const int thr = a.TW * threads_per_line(n, a.variant);
    const size_t smem = lds_elems(n, a.TW) * sizeof(cpx<T>);
    return dispatch_size<SET>(n, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        auto go = [&](auto VC) -> hipError_t {
            constexpr int VV = decltype(VC)::value;
            if constexpr (sizeof(T) == 4) {
                if (a.lpt == 2) return launch_k(k_pencils<T, NN, DIR, VV, 2>, dim3(blocks), dim3(thr), smem, s, a);
            }
            return launch_k(k_pencils<T, NN, DIR, VV, 1>, dim3(blocks), dim3(thr), smem, s, a);
        };
        if constexpr (plain_variant<NN>() == 0) {
            if (!a.variant) return go(std::integral_constant<int, 0>{});
        }
        return go(std::integral_constant<int, 1>{});
    });
}

// a grouped launch: the caller has set first/blocks/ncls, the per-class argument arrays
// (device memory, with each class's lines per block set for `threads` threads) and the
// dynamic LDS size smem, the largest any class needs
template <typename T>
hipError_t launch_rows_fused_group(const BatchGroupArgs<T>& g, int threads, size_t smem, hipStream_t s) {
    if (g.ncls < 1 || g.ncls > kMaxGroup || threads > kMaxBlock) return hipErrorInvalidValue;
    return launch_k(k_rows_fused_group<T>, dim3(g.first[g.ncls]), dim3(threads), smem, s, g);
}

template <typename T>
hipError_t launch_pencil_update_group(const BatchGroupArgs<T>& g, int threads, size_t smem, hipStream_t s) {
    if (g.ncls < 1 || g.ncls > kMaxGroup || threads > kMaxBlock) return hipErrorInvalidValue;
    return launch_k(k_pencil_update_group<T>, dim3(g.first[g.ncls]), dim3(threads), smem, s, g);
}


template <typename T>
hipError_t launch_bluestein(const BluesteinArgs<T>& b0, int n, hipStream_t s) {
    BluesteinArgs<T> b = b0;
    b.p.n = n;
    b.p.lpt = 1;
    if (n < 2 || 2 * n - 1 > b.m) return hipErrorInvalidValue;
    const long blocks = pencil_blocks(b.p.seg, b.p.nseg, b.p.TW, &b.p.tiles0);
    const size_t smem = lds_elems(b.m, b.p.TW) * sizeof(cpx<T>);
    auto go = [&](auto MC) -> hipError_t {
        constexpr int MM = decltype(MC)::value;
        constexpr int V = plain_variant<MM>();
        const int thr = b.p.TW * LineShape<MM, V>::T;
        return launch_k(k_bluestein<T, MM>, dim3(blocks), dim3(thr), smem, s, b);
    };
    switch (b.m) {
        case 128: return go(std::integral_constant<int, 128>{});
        case 256: return go(std::integral_constant<int, 256>{});
        case 512: return go(std::integral_constant<int, 512>{});
        case 1024: return go(std::integral_constant<int, 1024>{});
        case 2048: return go(std::integral_constant<int, 2048>{});
        default: return hipErrorInvalidValue;
    }
}

template <typename T, int SET>
hipError_t launch_pencil_update_set(const PencilUpdateArgs<T>& u0, int n, hipStream_t s) {
    PencilUpdateArgs<T> u = u0;
    u.p.n = n;
    u.p.lpt = 1;
    const long blocks = pencil_blocks(u.p.seg, u.p.nseg, u.p.TW, &u.p.tiles0);
    const int thr = u.p.TW * threads_per_line(n, 1);
    size_t smem = lds_elems(n, u.p.TW) * sizeof(cpx<T>);
    if (u.egrad_partial) smem += (size_t)thr * sizeof(double);
    return dispatch_size<SET>(n, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        if (u.egrad_partial) return launch_k(k_pencil_update<T, NN, true>, dim3(blocks), dim3(thr), smem, s, u);
        return launch_k(k_pencil_update<T, NN, false>, dim3(blocks), dim3(thr), smem, s, u);
    });
}

template <typename T, int SET>
hipError_t launch_rows_r2c_set(const RowR2CArgs<T>& a, hipStream_t s) {
    const long blocks = rows_blocks(a.rows, a.TW);
    const int thr = a.TW * threads_per_line(a.nx);
    // exchange lines and the block's copy of the twiddle table
    const size_t smem = (lds_elems(a.nx, a.TW) + a.nx) * sizeof(cpx<T>);
    return dispatch_size<SET>(a.nx, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        return launch_k(k_rows_r2c<T, NN>, dim3(blocks), dim3(thr), smem, s, a);
    });
}

template <typename T, int SET>
hipError_t launch_rows_c2r_set(const RowC2RArgs<T>& a, hipStream_t s) {
    const long blocks = rows_blocks(a.rows, a.TW);
    const int thr = a.TW * threads_per_line(a.nx);
    // exchange lines and the block's copy of the twiddle table
    const size_t smem = (lds_elems(a.nx, a.TW) + a.nx) * sizeof(cpx<T>);
    return dispatch_size<SET>(a.nx, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        return launch_k(k_rows_c2r<T, NN>, dim3(blocks), dim3(thr), smem, s, a);
    });
}

template <typename T, int SET>
hipError_t launch_rows_sync_set(const RowSyncArgs<T>& a, hipStream_t s) {
    if (a.nd < 1 || a.nd > kMaxArr || !a.fmain[0]) return hipErrorInvalidValue;
    if (a.mode == 1 && (!a.phi || a.nd != 1 || a.fmain[1])) return hipErrorInvalidValue;
    if (a.mode == 2 && (!a.fmain[1] || !a.fout[0] || a.nd != 1 || a.npoly < 0 || a.npoly > kMaxPoly))
        return hipErrorInvalidValue;
    const long blocks = rows_blocks(a.rows, a.TW);
    const int thr = a.TW * threads_per_line(a.nx);
    // exchange lines and the block's copy of the twiddle table
    const size_t smem = (lds_elems(a.nx, a.TW) + a.nx) * sizeof(cpx<T>);
    const int mode = (a.mode == 0) ? (a.nd <= 2 ? 0 : 3) : a.mode;
    return dispatch_size<SET>(a.nx, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        switch (mode) {
            case 0: return launch_k(k_rows_sync<T, NN, 0>, dim3(blocks), dim3(thr), smem, s, a);
            case 1: return launch_k(k_rows_sync<T, NN, 1>, dim3(blocks), dim3(thr), smem, s, a);
            case 2: return launch_k(k_rows_sync<T, NN, 2>, dim3(blocks), dim3(thr), smem, s, a);
            default: return launch_k(k_rows_sync<T, NN, 3>, dim3(blocks), dim3(thr), smem, s, a);
        }
    });
}

template <typename T, int SET>
hipError_t launch_rows_fused_set(const RowFusedArgs<T>& a, hipStream_t s) {
    const long blocks = rows_fused_blocks(a.rows, a.TW);
    const int thr = a.TW * threads_per_line(a.nx, a.variant);
    // exchange lines, the block's copy of the twiddle table, the monitor's partial sums
    size_t smem = (lds_elems(a.nx, a.TW) + a.nx) * sizeof(cpx<T>);
    if (a.bulk_partial) smem += (size_t)thr * sizeof(double);
    return dispatch_size<SET>(a.nx, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        if constexpr (plain_variant<NN>() == 0) {
            if (!a.variant) {
                if (a.bulk_partial) return launch_k(k_rows_fused<T, NN, 0, true>, dim3(blocks), dim3(thr), smem, s, a);
                return launch_k(k_rows_fused<T, NN, 0, false>, dim3(blocks), dim3(thr), smem, s, a);
            }
        }
        if (a.bulk_partial) return launch_k(k_rows_fused<T, NN, 1, true>, dim3(blocks), dim3(thr), smem, s, a);
        return launch_k(k_rows_fused<T, NN, 1, false>, dim3(blocks), dim3(thr), smem, s, a);
    });
}

// the public launch functions: extended lengths go to the set-1 instantiations, which live
// in units of their own (spectral_ext*.hip)
#define GPF_EXT_DECL(T)                                                                                  \
    extern template hipError_t launch_pencils_dir_set<T, -1, 1>(const PencilArgs<T>&, int, hipStream_t);  \
    extern template hipError_t launch_pencils_dir_set<T, 1, 1>(const PencilArgs<T>&, int, hipStream_t);   \
    extern template hipError_t launch_pencil_update_set<T, 1>(const PencilUpdateArgs<T>&, int, hipStream_t); \
    extern template hipError_t launch_rows_r2c_set<T, 1>(const RowR2CArgs<T>&, hipStream_t);             \
    extern template hipError_t launch_rows_c2r_set<T, 1>(const RowC2RArgs<T>&, hipStream_t);             \
    extern template hipError_t launch_rows_sync_set<T, 1>(const RowSyncArgs<T>&, hipStream_t);           \
    extern template hipError_t launch_rows_fused_set<T, 1>(const RowFusedArgs<T>&, hipStream_t);
GPF_EXT_DECL(float)
GPF_EXT_DECL(double)
#undef GPF_EXT_DECL

template <typename T, int DIR>
hipError_t launch_pencils_dir(const PencilArgs<T>& a, int n, hipStream_t s) {
    return ext_length(n) ? launch_pencils_dir_set<T, DIR, 1>(a, n, s) : launch_pencils_dir_set<T, DIR, 0>(a, n, s);
}

template <typename T>
hipError_t launch_pencil_update(const PencilUpdateArgs<T>& u, int n, hipStream_t s) {
    return ext_length(n) ? launch_pencil_update_set<T, 1>(u, n, s) : launch_pencil_update_set<T, 0>(u, n, s);
}

template <typename T>
hipError_t launch_rows_r2c(const RowR2CArgs<T>& a, hipStream_t s) {
    return ext_length(a.nx) ? launch_rows_r2c_set<T, 1>(a, s) : launch_rows_r2c_set<T, 0>(a, s);
}

template <typename T>
hipError_t launch_rows_c2r(const RowC2RArgs<T>& a, hipStream_t s) {
    return ext_length(a.nx) ? launch_rows_c2r_set<T, 1>(a, s) : launch_rows_c2r_set<T, 0>(a, s);
}

template <typename T>
hipError_t launch_rows_sync(const RowSyncArgs<T>& a, hipStream_t s) {
    return ext_length(a.nx) ? launch_rows_sync_set<T, 1>(a, s) : launch_rows_sync_set<T, 0>(a, s);
}

template <typename T>
hipError_t launch_rows_fused(const RowFusedArgs<T>& a, hipStream_t s) {
    return ext_length(a.nx) ? launch_rows_fused_set<T, 1>(a, s) : launch_rows_fused_set<T, 0>(a, s);
}

}  // namespace gpf
