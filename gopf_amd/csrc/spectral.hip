// spectral.hip -- hand-written gfx950 kernels of gopf's spectral time step.
//
// Every kernel is one HBM pass: a tile of lines is staged into LDS with
// coalesced loads, transformed in LDS (fft_core.h), optionally combined with
// pointwise / k-space work, and written back with coalesced stores.
//
// Kernels
//   k_pencils        C2C transform of pencils along a strided or contiguous axis
//                    (y / z passes over the half spectrum, and every axis of the
//                    full-complex FourierTransform drop-in, fftWrap.go:26-39)
//   k_pencil_update  forward transform -> semi-implicit k-space update -> inverse
//                    transform of the last axis (euler.go:27-39 with
//                    LaplacianN.Eval diffOp.go:25-29, ApplyModalFilter util.go:125-132)
//   k_rows_r2c       x-row transforms of pairs of real arrays into half spectra
//   k_rows_c2r       x-row inverse of half spectra into real rows (euler.go:42-45)
//   k_rows_fused     x-row inverse -> pointwise explicit source g(c) (SyncDerivedFields,
//                    model.go:237-241) -> x-row forward of (c, g); optional FP64
//                    energy/mass partial sums
#include <type_traits>

#include "fft_core.h"
#include "spectral.h"

namespace gpf {

namespace {

constexpr double kPi = 3.14159265358979323846;

template <typename T, int N, int DIR>
__device__ __forceinline__ void line_fft(cpx<T>* buf, int es, int n, int t, const cpx<T>* __restrict__ tw) {
    if constexpr (N > 0) {
        fft_line<T, N, DIR>(buf, es, t, true, tw);
    } else {
        dft_line_rt<T, DIR>(buf, es, n, t, true, tw);
    }
}

__device__ __forceinline__ double wrapf(int k, int n) {
    double f = (double)k / (double)n;
    return (f > 0.5) ? f - 1.0 : f;
}

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

// block-wide FP64 sum; red must hold blockDim.x doubles; result valid in thread 0
__device__ double block_sum(double v, double* red) {
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    red[tid] = v;
    __syncthreads();
    int p2 = 1;
    while (p2 < nt) p2 <<= 1;
    for (int s = p2 / 2; s > 0; s >>= 1) {
        if (tid < s && tid + s < nt) red[tid] += red[tid + s];
        __syncthreads();
    }
    double r = red[0];
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

// ---------------------------------------------------------------------------
// tile decode shared by the pencil kernels
// ---------------------------------------------------------------------------
struct TileInfo {
    int s;        // segment
    int ia, b0;   // a index and first b of the tile
    long base;    // element offset of line (ia, b0)
    int nvalid;   // valid lines in the tile
};

template <typename T>
__device__ __forceinline__ TileInfo decode_tile(const PencilArgs<T>& a) {
    TileInfo ti;
    int tile = blockIdx.x;
    ti.s = 0;
    if (tile >= a.tiles0) {
        tile -= a.tiles0;
        ti.s = 1;
    }
    const Seg& g = a.seg[ti.s];
    const int tpa = (g.nb + a.TW - 1) / a.TW;
    ti.ia = tile / tpa;
    ti.b0 = (tile - ti.ia * tpa) * a.TW;
    ti.base = (long)ti.ia * g.sa + (long)ti.b0 * g.sb;
    ti.nvalid = min(a.TW, g.nb - ti.b0);
    return ti;
}

// global tile <-> LDS (layout lds[p * TWP + l])
template <typename T>
__device__ __forceinline__ void tile_load(const cpx<T>* __restrict__ src, const Seg& g, const TileInfo& ti,
                                          cpx<T>* L, int n, int TW, int TWP) {
    const int nthr = blockDim.x;
    const int total = TW * n;
    if (g.es == 1) {
        for (int idx = threadIdx.x; idx < total; idx += nthr) {
            const int l = idx / n, p = idx - (idx / n) * n;
            cpx<T> v = mk<T>(T(0), T(0));
            if (l < ti.nvalid) v = src[ti.base + (long)l * g.sb + p];
            L[p * TWP + l] = v;
        }
    } else {
        for (int idx = threadIdx.x; idx < total; idx += nthr) {
            const int l = idx & (TW - 1), p = idx / TW;
            cpx<T> v = mk<T>(T(0), T(0));
            if (l < ti.nvalid) v = src[ti.base + (long)l * g.sb + (long)p * g.es];
            L[p * TWP + l] = v;
        }
    }
}

template <typename T>
__device__ __forceinline__ void tile_store(cpx<T>* __restrict__ dst, const Seg& g, const TileInfo& ti,
                                           const cpx<T>* L, int n, int TW, int TWP) {
    const int nthr = blockDim.x;
    const int total = TW * n;
    if (g.es == 1) {
        for (int idx = threadIdx.x; idx < total; idx += nthr) {
            const int l = idx / n, p = idx - (idx / n) * n;
            if (l < ti.nvalid) dst[ti.base + (long)l * g.sb + p] = L[p * TWP + l];
        }
    } else {
        for (int idx = threadIdx.x; idx < total; idx += nthr) {
            const int l = idx & (TW - 1), p = idx / TW;
            if (l < ti.nvalid) dst[ti.base + (long)l * g.sb + (long)p * g.es] = L[p * TWP + l];
        }
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <typename T, int N, int DIR>
__global__ void k_pencils(PencilArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    const int n = (N > 0) ? N : a.n;
    const int TW = a.TW, TWP = TW + 1;
    const TileInfo ti = decode_tile(a);
    const Seg& g = a.seg[ti.s];
    for (int q = 0; q < a.narr; ++q) {
        const cpx<T>* in = a.src[q][ti.s] ? a.src[q][ti.s] : a.arr[q][ti.s];
        tile_load(in, g, ti, lds + q * n * TWP, n, TW, TWP);
    }
    __syncthreads();
    const int l = threadIdx.x & (TW - 1), t = threadIdx.x / TW;
    for (int q = 0; q < a.narr; ++q) line_fft<T, N, DIR>(lds + q * n * TWP + l, TWP, n, t, a.tw);
    for (int q = 0; q < a.narr; ++q) tile_store(a.arr[q][ti.s], g, ti, lds + q * n * TWP, n, TW, TWP);
}

template <typename T, int N>
__global__ void k_pencil_update(PencilUpdateArgs<T> u) {
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    const PencilArgs<T>& a = u.p;
    const int n = (N > 0) ? N : a.n;
    const int TW = a.TW, TWP = TW + 1;
    const TileInfo ti = decode_tile(a);
    const Seg& g = a.seg[ti.s];
    cpx<T>* Lc = lds;
    cpx<T>* Lg = lds + n * TWP;
    tile_load(a.arr[0][ti.s], g, ti, Lc, n, TW, TWP);
    if (u.has_src) tile_load(a.arr[1][ti.s], g, ti, Lg, n, TW, TWP);
    __syncthreads();
    const int l = threadIdx.x & (TW - 1), t = threadIdx.x / TW;
    line_fft<T, N, -1>(Lc + l, TWP, n, t, a.tw);
    if (u.has_src) line_fft<T, N, -1>(Lg + l, TWP, n, t, a.tw);
    // k-space update of this tile
    double eg = 0.0;
    const int total = TW * n;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int ll = idx & (TW - 1), p = idx / TW;
        if (ll >= ti.nvalid) continue;
        const int src[3] = {ti.ia, ti.b0 + ll, p};
        int kc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) kc[d] = (g.ksrc[d] == 3) ? g.kfix[d] : src[g.ksrc[d]];
        const double fx = wrapf(kc[0], u.g.nx);
        const double fy = wrapf(kc[1], u.g.ny);
        const double fz = (u.g.ndim == 3) ? wrapf(kc[2], u.g.nz) : 0.0;
        const double f2 = fx * fx + fy * fy + fz * fz;
        const double L = -(4.0 * kPi * kPi) * f2;
        const double D = kpoly_eval(u.denom, L);
        const double E = kpoly_eval(u.lin, L);
        double A = 1.0 / (1.0 - u.dt * D);
        if (u.filter) A *= vandeven(u.filter, sqrt(f2) * 2.0 / kPi);
        double Lp = 1.0;
        for (int q = 0; q < u.src_lap; ++q) Lp *= L;
        const double m1 = A * (1.0 + u.dt * E);
        const double m2 = A * u.dt * Lp;
        cpx<T> c = Lc[p * TWP + ll];
        cpx<T> gs = u.has_src ? Lg[p * TWP + ll] : mk<T>(T(0), T(0));
        if (u.egrad_partial) {
            const double w = (kc[0] == 0 || 2 * kc[0] == u.g.nx) ? 1.0 : 2.0;
            eg += w * (-L) * ((double)c.x * (double)c.x + (double)c.y * (double)c.y);
        }
        cpx<T> r;
        if (u.mean && kc[0] == 0 && kc[1] == 0 && kc[2] == 0) {
            // k = 0 advances in FP64: the arrays hold c - mean, so their k = 0 coefficient stays 0
            // and mean_new = m1(0) mean + m2(0) G(0) / N
            const double mu = *u.mean;
            *u.mean = m1 * mu + m2 * (double)gs.x * u.scale;
            r = mk<T>(T(0), T(0));
        } else {
            const T s1 = (T)(m1 * u.scale), s2 = (T)(m2 * u.scale);
            r = mk<T>(s1 * c.x + s2 * gs.x, s1 * c.y + s2 * gs.y);
        }
        Lc[p * TWP + ll] = r;
    }
    if (u.egrad_partial) {
        double* red = reinterpret_cast<double*>(lds + (u.has_src ? 2 : 1) * n * TWP);
        __syncthreads();
        const double tot = block_sum(eg, red);
        if (threadIdx.x == 0) u.egrad_partial[blockIdx.x] = tot;
    }
    __syncthreads();
    line_fft<T, N, +1>(Lc + l, TWP, n, t, a.tw);
    tile_store(a.arr[0][ti.s], g, ti, Lc, n, TW, TWP);
}

template <typename T, int N>
__global__ void k_rows_r2c(RowR2CArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    const int n = (N > 0) ? N : a.nx;
    const int nh = n / 2;
    const int TW = a.TW, TWP = TW + 1;
    const long row0 = (long)blockIdx.x * TW;
    const int nvalid = (int)min((long)TW, a.rows - row0);
    const int npair = (a.narr + 1) / 2;
    for (int q2 = 0; q2 < npair; ++q2) {
        const T* ia = a.in[2 * q2];
        const T* ib = (2 * q2 + 1 < a.narr) ? a.in[2 * q2 + 1] : nullptr;
        cpx<T>* L = lds + q2 * n * TWP;
        for (int idx = threadIdx.x; idx < TW * n; idx += blockDim.x) {
            const int l = idx / n, p = idx - (idx / n) * n;
            T x = T(0), y = T(0);
            if (l < nvalid) {
                const long off = (row0 + l) * n + p;
                if (ia) x = ia[off] * a.in_scale;
                if (ib) y = ib[off] * a.in_scale;
            }
            L[p * TWP + l] = mk<T>(x, y);
        }
    }
    __syncthreads();
    const int l = threadIdx.x & (TW - 1), t = threadIdx.x / TW;
    for (int q2 = 0; q2 < npair; ++q2) line_fft<T, N, -1>(lds + q2 * n * TWP + l, TWP, n, t, a.tw);
    for (int q2 = 0; q2 < npair; ++q2) {
        const cpx<T>* L = lds + q2 * n * TWP;
        const int qa = 2 * q2, qb = 2 * q2 + 1;
        const int m = nh + 1;
        for (int idx = threadIdx.x; idx < TW * m; idx += blockDim.x) {
            const int ll = idx / m, k = idx - (idx / m) * m;
            if (ll >= nvalid) continue;
            const long row = row0 + ll;
            const cpx<T> z = L[k * TWP + ll];
            const cpx<T> zm = L[((n - k) % n) * TWP + ll];
            // A = (z + conj zm) / 2,  B = (z - conj zm) / (2 i)
            const cpx<T> A = mk<T>(T(0.5) * (z.x + zm.x), T(0.5) * (z.y - zm.y));
            const cpx<T> B = mk<T>(T(0.5) * (z.y + zm.y), T(-0.5) * (z.x - zm.x));
            if (k < nh) {
                a.main[qa][row * nh + k] = A;
                if (qb < a.narr) a.main[qb][row * nh + k] = B;
            } else {
                a.nyq[qa][row] = A;
                if (qb < a.narr) a.nyq[qb][row] = B;
            }
        }
    }
}

template <typename T, int N>
__global__ void k_rows_c2r(RowC2RArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
    const int n = (N > 0) ? N : a.nx;
    const int nh = n / 2;
    const int TW = a.TW, TWP = TW + 1;
    const long row0 = (long)blockIdx.x * TW;
    const int nvalid = (int)min((long)TW, a.rows - row0);
    const int npair = (a.narr + 1) / 2;
    const int m = nh + 1;
    for (int q2 = 0; q2 < npair; ++q2) {
        const int qa = 2 * q2, qb = 2 * q2 + 1;
        cpx<T>* L = lds + q2 * n * TWP;
        for (int idx = threadIdx.x; idx < TW * m; idx += blockDim.x) {
            const int ll = idx / m, k = idx - (idx / m) * m;
            cpx<T> A = mk<T>(T(0), T(0)), B = mk<T>(T(0), T(0));
            if (ll < nvalid) {
                const long row = row0 + ll;
                A = (k < nh) ? a.main[qa][row * nh + k] : a.nyq[qa][row];
                if (qb < a.narr) B = (k < nh) ? a.main[qb][row * nh + k] : a.nyq[qb][row];
            }
            // Z(k) = A(k) + i B(k); Z(n-k) = conj(A(k)) + i conj(B(k))
            L[k * TWP + ll] = mk<T>(A.x - B.y, A.y + B.x);
            if (k > 0 && k < nh) L[(n - k) * TWP + ll] = mk<T>(A.x + B.y, -A.y + B.x);
        }
    }
    __syncthreads();
    const int l = threadIdx.x & (TW - 1), t = threadIdx.x / TW;
    for (int q2 = 0; q2 < npair; ++q2) line_fft<T, N, +1>(lds + q2 * n * TWP + l, TWP, n, t, a.tw);
    for (int q2 = 0; q2 < npair; ++q2) {
        const int qa = 2 * q2, qb = 2 * q2 + 1;
        const cpx<T>* L = lds + q2 * n * TWP;
        for (int idx = threadIdx.x; idx < TW * n; idx += blockDim.x) {
            const int ll = idx / n, p = idx - (idx / n) * n;
            if (ll >= nvalid) continue;
            const long off = (row0 + ll) * n + p;
            const cpx<T> z = L[p * TWP + ll];
            a.out[qa][off] = z.x * a.scale;
            if (qb < a.narr) a.out[qb][off] = z.y * a.scale;
        }
    }
}

template <typename T, int N>
__global__ void k_rows_fused(RowFusedArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<T>* L = reinterpret_cast<cpx<T>*>(smem);
    const int n = (N > 0) ? N : a.nx;
    const int nh = n / 2;
    const int TW = a.TW, TWP = TW + 1;
    const long row0 = (long)blockIdx.x * TW;
    const int nvalid = (int)min((long)TW, a.rows - row0);
    const int m = nh + 1;
    // Hermitian extension of the state row
    for (int idx = threadIdx.x; idx < TW * m; idx += blockDim.x) {
        const int ll = idx / m, k = idx - (idx / m) * m;
        cpx<T> A = mk<T>(T(0), T(0));
        if (ll < nvalid) {
            const long row = row0 + ll;
            A = (k < nh) ? a.s_main[row * nh + k] : a.s_nyq[row];
        }
        L[k * TWP + ll] = A;
        if (k > 0 && k < nh) L[(n - k) * TWP + ll] = conjg(A);
    }
    __syncthreads();
    const int l = threadIdx.x & (TW - 1), t = threadIdx.x / TW;
    line_fft<T, N, +1>(L + l, TWP, n, t, a.tw);
    // real space: c, g(c), optional output and FP64 partial sums
    KSum sb, sm;
    const double mu = a.mean ? *a.mean : 0.0;
    const T muT = (T)mu;
    for (int idx = threadIdx.x; idx < TW * n; idx += blockDim.x) {
        const int ll = idx / n, p = idx - (idx / n) * n;
        const T dc = L[p * TWP + ll].x;   // c - mean (mean = 0 when the arrays carry k = 0)
        const T c = dc + muT;
        T gv = T(0);
        for (int q = a.npoly - 1; q >= 0; --q) gv = gv * c + (T)a.gpoly[q];
        L[p * TWP + ll] = mk<T>(dc, gv);
        if (ll < nvalid) {
            if (a.c_out) a.c_out[(row0 + ll) * n + p] = c;
            if (a.bulk_partial) {
                const double cd = (double)dc + mu;
                double f = 0.0;
                for (int q = a.nbulk - 1; q >= 0; --q) f = f * cd + a.bulk[q];
                sb.add(f);
                sm.add(cd);
            }
        }
    }
    if (a.bulk_partial) {
        double* red = reinterpret_cast<double*>(L + n * TWP);
        __syncthreads();
        const double b = block_sum(sb.get(), red);
        const double ms = block_sum(sm.get(), red);
        if (threadIdx.x == 0) {
            a.bulk_partial[blockIdx.x] = b;
            a.mass_partial[blockIdx.x] = ms;
        }
    }
    __syncthreads();
    line_fft<T, N, -1>(L + l, TWP, n, t, a.tw);
    for (int idx = threadIdx.x; idx < TW * m; idx += blockDim.x) {
        const int ll = idx / m, k = idx - (idx / m) * m;
        if (ll >= nvalid) continue;
        const long row = row0 + ll;
        const cpx<T> z = L[k * TWP + ll];
        const cpx<T> zm = L[((n - k) % n) * TWP + ll];
        const cpx<T> C = mk<T>(T(0.5) * (z.x + zm.x), T(0.5) * (z.y - zm.y));
        const cpx<T> G = mk<T>(T(0.5) * (z.y + zm.y), T(-0.5) * (z.x - zm.x));
        if (k < nh) {
            a.c_main[row * nh + k] = C;
            a.g_main[row * nh + k] = G;
        } else {
            a.c_nyq[row] = C;
            a.g_nyq[row] = G;
        }
    }
}

int threads_per_line(int n) {
    switch (n) {
#define GPF_T(NN) case NN: return Sched<NN>::T;
        GPF_T(2) GPF_T(4) GPF_T(8) GPF_T(16) GPF_T(32) GPF_T(64) GPF_T(128) GPF_T(256) GPF_T(512)
        GPF_T(1024) GPF_T(12) GPF_T(24) GPF_T(48) GPF_T(96) GPF_T(192) GPF_T(384) GPF_T(768)
#undef GPF_T
        default: return n;  // brute-force DFT: one thread per output
    }
}

template <typename F>
hipError_t dispatch_size(int n, F&& f) {
    switch (n) {
#define GPF_D(NN) case NN: return f(std::integral_constant<int, NN>{});
        GPF_D(2) GPF_D(4) GPF_D(8) GPF_D(16) GPF_D(32) GPF_D(64) GPF_D(128) GPF_D(256) GPF_D(512)
        GPF_D(1024) GPF_D(12) GPF_D(24) GPF_D(48) GPF_D(96) GPF_D(192) GPF_D(384) GPF_D(768)
#undef GPF_D
        default:
            if (n >= 1 && n <= 64) return f(std::integral_constant<int, 0>{});
            return hipErrorInvalidValue;
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

bool fft_size_fast(int n) {
    switch (n) {
        case 2: case 4: case 8: case 16: case 32: case 64: case 128: case 256: case 512: case 1024:
        case 12: case 24: case 48: case 96: case 192: case 384: case 768:
            return true;
        default:
            return false;
    }
}

bool fft_size_supported(int n) { return fft_size_fast(n) || (n >= 1 && n <= 64); }

int fft_threads_per_line(int n) { return threads_per_line(n); }

long rows_blocks(long rows, int TW) { return (rows + TW - 1) / TW; }

long pencil_blocks(const Seg* seg, int nseg, int TW, int* tiles0) {
    long tot = 0;
    for (int s = 0; s < nseg; ++s) {
        long t = (long)seg[s].na * ((seg[s].nb + TW - 1) / TW);
        if (s == 0 && tiles0) *tiles0 = (int)t;
        tot += t;
    }
    return tot;
}

template <typename T>
hipError_t launch_pencils(const PencilArgs<T>& a0, int n, int dir, hipStream_t s) {
    PencilArgs<T> a = a0;
    a.n = n;
    const long blocks = pencil_blocks(a.seg, a.nseg, a.TW, &a.tiles0);
    const int thr = a.TW * threads_per_line(n);
    const size_t smem = (size_t)a.narr * n * (a.TW + 1) * sizeof(cpx<T>);
    return dispatch_size(n, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        if (dir < 0) return launch_k(k_pencils<T, NN, -1>, dim3(blocks), dim3(thr), smem, s, a);
        return launch_k(k_pencils<T, NN, +1>, dim3(blocks), dim3(thr), smem, s, a);
    });
}

template <typename T>
hipError_t launch_pencil_update(const PencilUpdateArgs<T>& u0, int n, hipStream_t s) {
    PencilUpdateArgs<T> u = u0;
    u.p.n = n;
    const long blocks = pencil_blocks(u.p.seg, u.p.nseg, u.p.TW, &u.p.tiles0);
    const int thr = u.p.TW * threads_per_line(n);
    size_t smem = (size_t)(u.has_src ? 2 : 1) * n * (u.p.TW + 1) * sizeof(cpx<T>);
    if (u.egrad_partial) smem += (size_t)thr * sizeof(double);
    return dispatch_size(n, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        return launch_k(k_pencil_update<T, NN>, dim3(blocks), dim3(thr), smem, s, u);
    });
}

template <typename T>
hipError_t launch_rows_r2c(const RowR2CArgs<T>& a, hipStream_t s) {
    const long blocks = rows_blocks(a.rows, a.TW);
    const int thr = a.TW * threads_per_line(a.nx);
    const size_t smem = (size_t)((a.narr + 1) / 2) * a.nx * (a.TW + 1) * sizeof(cpx<T>);
    return dispatch_size(a.nx, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        return launch_k(k_rows_r2c<T, NN>, dim3(blocks), dim3(thr), smem, s, a);
    });
}

template <typename T>
hipError_t launch_rows_c2r(const RowC2RArgs<T>& a, hipStream_t s) {
    const long blocks = rows_blocks(a.rows, a.TW);
    const int thr = a.TW * threads_per_line(a.nx);
    const size_t smem = (size_t)((a.narr + 1) / 2) * a.nx * (a.TW + 1) * sizeof(cpx<T>);
    return dispatch_size(a.nx, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        return launch_k(k_rows_c2r<T, NN>, dim3(blocks), dim3(thr), smem, s, a);
    });
}

template <typename T>
hipError_t launch_rows_fused(const RowFusedArgs<T>& a, hipStream_t s) {
    const long blocks = rows_blocks(a.rows, a.TW);
    const int thr = a.TW * threads_per_line(a.nx);
    size_t smem = (size_t)a.nx * (a.TW + 1) * sizeof(cpx<T>);
    if (a.bulk_partial) smem += (size_t)thr * sizeof(double);
    return dispatch_size(a.nx, [&](auto NC) -> hipError_t {
        constexpr int NN = decltype(NC)::value;
        return launch_k(k_rows_fused<T, NN>, dim3(blocks), dim3(thr), smem, s, a);
    });
}

#define GPF_INST(T)                                                                                   \
    template hipError_t launch_pencils<T>(const PencilArgs<T>&, int, int, hipStream_t);             \
    template hipError_t launch_pencil_update<T>(const PencilUpdateArgs<T>&, int, hipStream_t);       \
    template hipError_t launch_rows_r2c<T>(const RowR2CArgs<T>&, hipStream_t);                      \
    template hipError_t launch_rows_c2r<T>(const RowC2RArgs<T>&, hipStream_t);                      \
    template hipError_t launch_rows_fused<T>(const RowFusedArgs<T>&, hipStream_t);
GPF_INST(float)
GPF_INST(double)
#undef GPF_INST

}  // namespace gpf
