// pointwise.hip -- real-space and k-space pointwise kernels of the general stepper.
//
//   k_derived     SyncDerivedFields / DerivedFieldCalcFromDesc (model.go:237-241,
//                 util.go:38-65) and RegisterFunction polynomials (model.go:400-412)
//   k_kupdate     GetRHS / GetDenum sums of ConcreteTerm closures (rhsBuilder.go:125-190,
//                 model.go:283-314) fused with the Euler / RK4 stage algebra
//                 (euler.go:32-38, rk4.go:77-127) and ApplyModalFilter (util.go:125-132)
//   k_etd_residual  the spectral part of ImplicitEuler's residual (implicitEuler.go:68-94)
//   k_kgrad       i 2 pi f_d u(k) with f_d = +0.5 zeroed (squareGradientTerm.go:44-50)
//   k_kelastic    scalar Khachaturyan kernel K(k): force (effectiveForce.go:26-35), 3x3
//                 displacement solve (linearElasticity.go:33-63), strains
//                 (linearElasticity.go:67-83) and the sigma*:eps sum (homoLinElast.go:72-89)
//                 collapsed by linearity into one multiplier of FFT(H)
#include "blockred.h"
#include "fft_core.h"
#include "pointwise.h"
#include "kspace.h"

namespace gpf {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kStageThreads = 256;  // threads per block of the k-space stage kernel
constexpr int kStageEpt = 4;        // its elements per thread

__device__ __forceinline__ double vandeven(const double* __restrict__ tab, double x) {
    const double N = 999.0;
    int idx = (int)(x * N);
    if (idx >= 999) return tab[999];
    double dx = 1.0 / N;
    double dy = tab[idx + 1] - tab[idx];
    double x0 = (double)idx * dx;
    return tab[idx] + (x - x0) * dy / dx;
}

// math.Pow(x, n) for a small non-negative integer n: Go multiplies the (renormalised)
// mantissas by repeated squaring, which rounds as this loop does
__device__ __forceinline__ double go_ipow(double x, int n) {
    double a = 1.0, b = x;
    for (int i = n; i != 0; i >>= 1) {
        if (i & 1) a *= b;
        b *= b;
    }
    return a;
}

// interpolant (spectralViscosity.go:29-39), including its value -1 at f = peak
__device__ __forceinline__ double sv_interp(double f, double peak) {
    const double frac = 1.0 / 3.0;
    if (f < frac * peak) return 0.0;
    if (f > peak) return 1.0;
    const double x = 1.5 * (f - frac * peak) / peak;
    return 2.0 * x * x - 3.0 * x * x * x;
}

// D(k) = sum of the LAP^p, SpectralViscosity and TensorialHessian terms (f2 = |f|^2; f has
// ndim components in the reference's Freq order)
__device__ __forceinline__ double denominator(const KPoly& p, const SvPoly& sv, const HessPoly& hs, int ndim,
                                              const double f[3], double L, double f2) {
    double D = 0.0;
    for (int t = 0; t < p.n; ++t) D += p.coef[t] * ipow(L, p.pow[t]);
    if (sv.n) {
        const double fr = sqrt(f2);
        for (int t = 0; t < sv.n; ++t)
            D += sv.coef[t] * ipow(L, sv.lap[t]) * sv_interp(fr, sv.thr[t]) * go_ipow(fr, sv.pw[t]);
    }
    // the loops over axes have fixed trip counts (ndim enters as a guard), so that f stays in
    // registers
    for (int t = 0; t < hs.n; ++t) {
        const int d = hs.kd[t];
        double h = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            if (j < ndim) h += -4.0 * kPi * kPi * f[j] * f[j] * hs.K[t][j * d + j];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = j + 1; k < 3; ++k)
                if (k < ndim) h += -8.0 * kPi * kPi * f[j] * f[k] * hs.K[t][j * d + k];
        D += hs.coef[t] * ipow(L, hs.lap[t]) * h;
    }
    return D;
}

// The reference's Freq gives the Nyquist index +0.5 for k and for its conjugate -k, so an
// off-diagonal f_j f_k with exactly one Nyquist component changes sign between the two and
// D(k) != D(-k) there: the reference's update is not Hermitian and its fields pick up an
// imaginary part.  The device keeps real fields, i.e. the Hermitian part of the reference's
// update, which is the mean of the update evaluated with D(k) and with D(-k).  conj_freq
// gives the Freq vector of -k; asym_row tells whether D can differ at k.
__device__ __forceinline__ bool asym_row(const HessPoly& hs, int ndim, const double f[3]) {
    if (hs.n == 0) return false;
    bool any = false, all = true;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        if (j >= ndim) continue;
        const bool nyq = (f[j] == 0.5);
        any |= nyq;
        all &= nyq;
    }
    return any && !all;
}

__device__ __forceinline__ void conj_freq(int ndim, const double f[3], double fc[3]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) fc[j] = (j < ndim && f[j] != 0.5) ? -f[j] : f[j];
}

template <typename T>
__device__ __forceinline__ T powT(T x, double p) {
    const double r = rint(p);
    if (r == p && fabs(p) <= 64.0) {
        int ip = (int)fabs(p);
        T v = T(1);
        T b = x;
        while (ip) {
            if (ip & 1) v *= b;
            b *= b;
            ip >>= 1;
        }
        return (p < 0) ? T(1) / v : v;
    }
    return (T)pow((double)x, p);
}

template <typename T>
__global__ void k_derived(DerivedArgs<T> a) {
    if (a.noise) {
        for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.n; i += (long)gridDim.x * blockDim.x)
            a.out[i] = (T)(a.noise_std * NoiseDraw::normal(i, a.counter, a.stream, a.seed));
        return;
    }
    if (a.dense) {
        // Horner in f0 of polynomials in f1
        for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.n; i += (long)gridDim.x * blockDim.x) {
            const T x0 = a.f[0][i];
            const T x1 = (a.nf > 1) ? a.f[1][i] : T(0);
            T acc = T(0);
#pragma unroll
            for (int p = 3; p >= 0; --p) {
                const T inner = (((T)a.dc[p][3] * x1 + (T)a.dc[p][2]) * x1 + (T)a.dc[p][1]) * x1 + (T)a.dc[p][0];
                acc = acc * x0 + inner;
            }
            a.out[i] = acc;
        }
        return;
    }
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.n; i += (long)gridDim.x * blockDim.x) {
        T x[kMaxFields];
        for (int f = 0; f < a.nf; ++f) x[f] = a.f[f] ? a.f[f][i] : T(0);
        T acc = T(0);
        for (int t = 0; t < a.nterm; ++t) {
            T v = (T)a.coef[t];
            for (int f = 0; f < a.nf; ++f)
                if (a.pw[t][f] != 0.0) v *= powT(x[f], a.pw[t][f]);
            acc += v;
        }
        a.out[i] = acc;
    }
}

template <typename T>
__global__ void k_kupdate(KUpdateArgs<T> a) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        int kx, ky, kz;
        kcoords(a.g, i, a.nmain, kx, ky, kz);
        double f[3];
        ref_freq(a.g, kx, ky, kz, f);
        const double f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
        const double L = -(4.0 * kPi * kPi) * f2;
        double rx = 0.0, ry = 0.0;
        for (int t = 0; t < a.nterm; ++t) {
            double m = a.coef[t] * ipow(L, a.pw[t]);
            if (a.dmul[t]) m *= a.dmul[t][0];
            if (a.tmode[t] == 1) {
                const double fc = f[a.tcomp[t]];
                if (fabs(fabs(fc) - 0.5) > 1e-6) {
                    const double w = m * 2.0 * sin(kPi * fc);
                    const cpx<T> b = a.brick[t][i];
                    rx -= w * (double)b.y;
                    ry += w * (double)b.x;
                }
                continue;
            }
            if (a.brick[t]) {
                const cpx<T> b = a.brick[t][i];
                rx += m * (double)b.x;
                ry += m * (double)b.y;
            } else {
                rx += m;
            }
        }
        const double D = denominator(a.denom, a.sv, a.hs, a.g.ndim, f, L, f2);
        double s = 1.0 / (1.0 - a.gamma * D);
        if (asym_row(a.hs, a.g.ndim, f)) {
            double fc[3];
            conj_freq(a.g.ndim, f, fc);
            const double Dc = denominator(a.denom, a.sv, a.hs, a.g.ndim, fc, L, f2);
            s = 0.5 * (s + 1.0 / (1.0 - a.gamma * Dc));
        }
        if (a.filter) s *= vandeven(a.filter, sqrt(f2) * 2.0 / kPi);
        double cx = 0.0, cy = 0.0;
        if (a.cur) {
            const cpx<T> c = a.cur[i];
            cx = (double)c.x;
            cy = (double)c.y;
        }
        a.dst[i] = mk<T>((T)(s * (a.alpha * cx + a.beta * rx)), (T)(s * (a.alpha * cy + a.beta * ry)));
    }
}

// RK4.Step's stage algebra (rk4.go:35-127) fused with GetRHS; see Rk4StageArgs.  A thread
// takes EPT elements kStageThreads apart: the uniform data of each term is read once for
// EPT elements, and the loads of all EPT elements are in flight together.
template <typename T, int EPT>
__global__ void __launch_bounds__(kStageThreads) k_rk4_stage(Rk4StageArgs<T> a) {
    const long i0 = (long)blockIdx.x * (kStageThreads * EPT) + threadIdx.x;
    double f[EPT][3], f2[EPT], L[EPT], rx[EPT], ry[EPT];
    bool ok[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const long i = i0 + kStageThreads * e;
        ok[e] = i < a.nk;
        int kx = 0, ky = 0, kz = 0;
        if (ok[e]) kcoords(a.g, i, a.nmain, kx, ky, kz);
        ref_freq(a.g, kx, ky, kz, f[e]);
        f2[e] = f[e][0] * f[e][0] + f[e][1] * f[e][1] + f[e][2] * f[e][2];
        L[e] = -(4.0 * kPi * kPi) * f2[e];
        rx[e] = 0.0;
        ry[e] = 0.0;
    }
    for (int t = 0; t < a.nterm; ++t) {
        const double c = a.coef[t];
        const int pw = a.pw[t];
        const double* dmp = a.dmul[t];
        const double dm = dmp ? dmp[0] : 1.0;
        const cpx<T>* br = a.brick[t];
        const int mode = a.tmode[t], comp = a.tcomp[t];
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            if (!ok[e]) continue;
            const long i = i0 + kStageThreads * e;
            double m = c * ipow(L[e], pw);
            if (dmp) m *= dm;
            if (mode == 1) {
                const double fc = (comp == 0) ? f[e][0] : (comp == 1) ? f[e][1] : f[e][2];
                if (fabs(fabs(fc) - 0.5) > 1e-6) {
                    const double w = m * 2.0 * sin(kPi * fc);
                    const cpx<T> b = br[i];
                    rx[e] -= w * (double)b.y;
                    ry[e] += w * (double)b.x;
                }
                continue;
            }
            if (br) {
                const cpx<T> b = br[i];
                rx[e] += m * (double)b.x;
                ry[e] += m * (double)b.y;
            } else {
                rx[e] += m;
            }
        }
    }
    const double gam = (a.stage == 3) ? a.dt : a.fc * a.dt;
    const double bp = a.fp * a.dt;
    const double bc = a.fc * a.dt;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        if (!ok[e]) continue;
        const long i = i0 + kStageThreads * e;
        // KF is stored in T between the stage kernels of RK4.Step; round it the same way
        const double qx = (double)(T)rx[e];
        const double qy = (double)(T)ry[e];
        const double D = denominator(a.denom, a.sv, a.hs, a.g.ndim, f[e], L[e], f2[e]);
        double s = 1.0 / (1.0 - gam * D);
        if (asym_row(a.hs, a.g.ndim, f[e])) {
            double fcj[3];
            conj_freq(a.g.ndim, f[e], fcj);
            const double Dc = denominator(a.denom, a.sv, a.hs, a.g.ndim, fcj, L[e], f2[e]);
            s = 0.5 * (s + 1.0 / (1.0 - gam * Dc));
        }
        if (a.stage == 0) {
            const cpx<T> k = a.k[i];
            a.init[i] = k;
            a.final[i] = mk<T>((T)((double)k.x + bp * qx), (T)((double)k.y + bp * qy));
            a.kn[i] = mk<T>((T)(s * ((double)k.x + bc * qx)), (T)(s * ((double)k.y + bc * qy)));
        } else if (a.stage < 3) {
            const cpx<T> fin = a.final[i], ini = a.init[i];
            a.final[i] = mk<T>((T)((double)fin.x + bp * qx), (T)((double)fin.y + bp * qy));
            a.kn[i] = mk<T>((T)(s * ((double)ini.x + bc * qx)), (T)(s * ((double)ini.y + bc * qy)));
        } else {
            if (a.filter) s *= vandeven(a.filter, sqrt(f2[e]) * 2.0 / kPi);
            const cpx<T> fin = a.final[i];
            // final is stored in T before the division (rk4.go:57-67)
            const T fx = (T)((double)fin.x + bp * qx), fy = (T)((double)fin.y + bp * qy);
            a.kn[i] = mk<T>((T)(s * (double)fx), (T)(s * (double)fy));
        }
    }
}

// the per-plan factor table of one stage kind (see StageDenomArgs): the FP64 arithmetic of
// k_rk4_stage's factor, done once per plan instead of in every stage
template <typename T>
__global__ void k_stage_denoms(StageDenomArgs<T> a) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        int kx, ky, kz;
        kcoords(a.g, i, a.nmain, kx, ky, kz);
        double f[3];
        ref_freq(a.g, kx, ky, kz, f);
        const double f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
        const double L = -(4.0 * kPi * kPi) * f2;
        const double D = denominator(a.denom, a.sv, a.hs, a.g.ndim, f, L, f2);
        double s = 1.0 / (1.0 - a.gamma * D);
        if (asym_row(a.hs, a.g.ndim, f)) {
            double fcj[3];
            conj_freq(a.g.ndim, f, fcj);
            const double Dc = denominator(a.denom, a.sv, a.hs, a.g.ndim, fcj, L, f2);
            s = 0.5 * (s + 1.0 / (1.0 - a.gamma * Dc));
        }
        if (a.filter) s *= vandeven(a.filter, sqrt(f2) * 2.0 / kPi);
        a.out[i] = (T)s;
    }
}

// k_rk4_stage with the stage factor read from its table (every term of tmode 0).  A thread
// takes EPT elements kStageThreads apart and issues every load of all of them (the table,
// the stage operands and the first kPre bricks) before any arithmetic, so the kernel keeps
// many loads in flight; L(k) is formed only when some term has a LAP power (NEEDL).
template <typename T, int EPT, bool NEEDL>
__global__ void __launch_bounds__(kStageThreads) k_rk4_stage_tab(Rk4StageArgs<T> a) {
    constexpr int kPre = 3;
    const long i0 = (long)blockIdx.x * (kStageThreads * EPT) + threadIdx.x;
    long iv[EPT];
    bool ok[EPT];
    T sf[EPT];
    cpx<T> o1[EPT], o2[EPT], br[kPre][EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const long i = i0 + kStageThreads * e;
        ok[e] = i < a.nk;
        iv[e] = ok[e] ? i : 0;  // past the end: a valid element, never stored
        sf[e] = a.stab[iv[e]];
        o1[e] = (a.stage == 0) ? a.k[iv[e]] : a.final[iv[e]];
        o2[e] = (a.stage == 1 || a.stage == 2) ? a.init[iv[e]] : mk<T>(T(0), T(0));
#pragma unroll
        for (int t = 0; t < kPre; ++t)
            br[t][e] = (t < a.nterm && a.brick[t]) ? a.brick[t][iv[e]] : mk<T>(T(0), T(0));
    }
    double L[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        L[e] = 0.0;
        if constexpr (NEEDL) {
            int kx = 0, ky = 0, kz = 0;
            kcoords(a.g, iv[e], a.nmain, kx, ky, kz);
            double f[3];
            ref_freq(a.g, kx, ky, kz, f);
            L[e] = -(4.0 * kPi * kPi) * (f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        }
    }
    double rx[EPT], ry[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) rx[e] = ry[e] = 0.0;
    for (int t = 0; t < a.nterm; ++t) {
        const double c = a.coef[t];
        const double* dmp = a.dmul[t];
        const double dm = dmp ? dmp[0] : 1.0;
        const int pw = a.pw[t];
        const cpx<T>* bp = a.brick[t];
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            double me = NEEDL ? c * ipow(L[e], pw) : c;
            if (dmp) me *= dm;
            if (!bp) {
                rx[e] += me;
                continue;
            }
            cpx<T> b;
            if (t == 0) b = br[0][e];
            else if (t == 1) b = br[1][e];
            else if (t == 2) b = br[2][e];
            else b = bp[iv[e]];
            rx[e] += me * (double)b.x;
            ry[e] += me * (double)b.y;
        }
    }
    const double bp = a.fp * a.dt;
    const double bc = a.fc * a.dt;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        if (!ok[e]) continue;
        const long i = iv[e];
        // KF is stored in T between the stage kernels of RK4.Step; round it the same way
        const double qx = (double)(T)rx[e];
        const double qy = (double)(T)ry[e];
        const double s = (double)sf[e];
        if (a.stage == 0) {
            const cpx<T> k = o1[e];
            a.init[i] = k;
            a.final[i] = mk<T>((T)((double)k.x + bp * qx), (T)((double)k.y + bp * qy));
            a.kn[i] = mk<T>((T)(s * ((double)k.x + bc * qx)), (T)(s * ((double)k.y + bc * qy)));
        } else if (a.stage < 3) {
            const cpx<T> fin = o1[e], ini = o2[e];
            a.final[i] = mk<T>((T)((double)fin.x + bp * qx), (T)((double)fin.y + bp * qy));
            a.kn[i] = mk<T>((T)(s * ((double)ini.x + bc * qx)), (T)(s * ((double)ini.y + bc * qy)));
        } else {
            const cpx<T> fin = o1[e];
            // final is stored in T before the division (rk4.go:57-67)
            const T fx = (T)((double)fin.x + bp * qx), fy = (T)((double)fin.y + bp * qy);
            a.kn[i] = mk<T>((T)(s * (double)fx), (T)(s * (double)fy));
        }
    }
}

// e^{D dt} and nonlinearIntegral (implicitEuler.go:151-162) for real D
__device__ __forceinline__ void etd_update(double D, double dt, double rx, double ry, double px, double py, double& e,
                                           double& ix, double& iy) {
    e = exp(D * dt);
    if (fabs(D) < 1e-5) {
        ix = 0.5 * dt * (rx + px * e);
        iy = 0.5 * dt * (ry + py * e);
    } else {
        const double c1 = (e - 1.0) / D;
        const double c2 = (e - D * dt - 1.0) / (D * D);
        ix = px * c1 + (rx - px) / dt * c2;
        iy = py * c1 + (ry - py) / dt * c2;
    }
}

// ImplicitEuler.updateEquation's spectral part (implicitEuler.go:79-88) with
// nonlinearIntegral (implicitEuler.go:151-162); D is real for the device's implicit
// terms (scalar * LAP^p), so e^{D dt} is real
template <typename T>
__global__ void k_etd_residual(EtdArgs<T> a) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        int kx, ky, kz;
        kcoords(a.g, i, a.nmain, kx, ky, kz);
        double f[3];
        ref_freq(a.g, kx, ky, kz, f);
        const double f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
        const double L = -(4.0 * kPi * kPi) * f2;
        double rx = 0.0, ry = 0.0;
        for (int t = 0; t < a.nterm; ++t) {
            double m = a.coef[t] * ipow(L, a.pw[t]);
            if (a.dmul[t]) m *= a.dmul[t][0];
            if (a.tmode[t] == 1) {
                const double fc = f[a.tcomp[t]];
                if (fabs(fabs(fc) - 0.5) > 1e-6) {
                    const double w = m * 2.0 * sin(kPi * fc);
                    const cpx<T> b = a.brick[t][i];
                    rx -= w * (double)b.y;
                    ry += w * (double)b.x;
                }
                continue;
            }
            if (a.brick[t]) {
                const cpx<T> b = a.brick[t][i];
                rx += m * (double)b.x;
                ry += m * (double)b.y;
            } else {
                rx += m;
            }
        }
        const cpx<T> p = a.rprev[i];
        const double px = (double)p.x, py = (double)p.y;
        double e, ix, iy;
        etd_update(denominator(a.denom, a.sv, a.hs, a.g.ndim, f, L, f2), a.dt, rx, ry, px, py, e, ix, iy);
        if (asym_row(a.hs, a.g.ndim, f)) {
            double fc[3], ec, icx, icy;
            conj_freq(a.g.ndim, f, fc);
            etd_update(denominator(a.denom, a.sv, a.hs, a.g.ndim, fc, L, f2), a.dt, rx, ry, px, py, ec, icx, icy);
            e = 0.5 * (e + ec);
            ix = 0.5 * (ix + icx);
            iy = 0.5 * (iy + icy);
        }
        const cpx<T> o = a.orig[i];
        const cpx<T> c = a.cur[i];
        a.dst[i] = mk<T>((T)((double)c.x - ((double)o.x * e + ix)), (T)((double)c.y - ((double)o.y * e + iy)));
    }
}

template <typename T>
__global__ void k_volcons_update(const double* __restrict__ px, long nb, const T* __restrict__ ind, double dt,
                                 double* __restrict__ st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double fi = 0.0;
    for (long b = 0; b < nb; ++b) fi += px[b];
    const double ii = (double)ind[0];
    if (st[2] != 0.0) {
        st[1] = fi;
        st[2] = 0.0;
    } else {
        const double d = fi - st[1];
        st[1] = fi;
        st[0] = st[0] - d / (dt * ii);
    }
}

template <typename T>
__global__ void k_kgrad(KGradArgs<T> a, int ndim) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        int kx, ky, kz;
        kcoords(a.g, i, a.nmain, kx, ky, kz);
        double f[3];
        ref_freq(a.g, kx, ky, kz, f);
        const cpx<T> u = a.src[i];
        for (int d = 0; d < ndim; ++d) {
            double fd = f[d];
            if (fabs(fd - 0.5) < 1e-10) fd = 0.0;
            const double w = 2.0 * kPi * fd;
            // (0 + i w) * u
            a.out[d][i] = mk<T>((T)(-w * (double)u.y), (T)(w * (double)u.x));
        }
    }
}

// 3x3 LU with partial pivoting; returns false if singular.  Every array index is a
// compile-time constant (the pivot row is swapped in by selects), so A stays in registers.
__device__ __forceinline__ bool solve3(double A[3][3], double b[3], double x[3]) {
    int piv[3] = {0, 1, 2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int p = k;
        double best = fabs(A[k][k]);
#pragma unroll
        for (int i = k + 1; i < 3; ++i)
            if (fabs(A[i][k]) > best) { best = fabs(A[i][k]); p = i; }
        if (best == 0.0) return false;
#pragma unroll
        for (int i = k + 1; i < 3; ++i) {
            const bool sw = (i == p);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double t = A[k][j];
                A[k][j] = sw ? A[i][j] : t;
                A[i][j] = sw ? t : A[i][j];
            }
            const int t = piv[k];
            piv[k] = sw ? piv[i] : t;
            piv[i] = sw ? t : piv[i];
        }
#pragma unroll
        for (int i = k + 1; i < 3; ++i) {
            A[i][k] /= A[k][k];
#pragma unroll
            for (int j = k + 1; j < 3; ++j) A[i][j] -= A[i][k] * A[k][j];
        }
    }
    double y[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double s = (piv[i] == 0) ? b[0] : (piv[i] == 1) ? b[1] : b[2];
#pragma unroll
        for (int j = 0; j < i; ++j) s -= A[i][j] * y[j];
        y[i] = s;
    }
#pragma unroll
    for (int i = 2; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int j = i + 1; j < 3; ++j) s -= A[i][j] * x[j];
        x[i] = s / A[i][i];
    }
    return true;
}

__device__ double elastic_kernel(const ElasticParams& ep, const double f[3]) {
    if (fabs(f[0]) < 1e-10 && fabs(f[1]) < 1e-10 && fabs(f[2]) < 1e-10) return 0.0;
    // every loop has a fixed trip count (dim enters as a guard) so that the local arrays
    // stay in registers
    double G[3][3];
    const double tp2 = 4.0 * kPi * kPi;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            double e = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int l = 0; l < 3; ++l) e += ep.C[m * 27 + j * 9 + n * 3 + l] * f[j] * f[l];
            G[m][n] = e * tp2;
        }
    // force direction b_c = sum_j sigma*_cj f_j for c < Dim (u = -i 2 pi G^-1 b H)
    double b[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (c >= ep.dim) continue;
#pragma unroll
        for (int j = 0; j < 3; ++j) b[c] += ep.sig[c * 3 + j] * f[j];
    }
    double v[3];
    if (!solve3(G, b, v)) return 0.0;
    double fm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) fm[c] = (fabs(fabs(f[c]) - 0.5) < 1e-10) ? 0.0 : f[c];
    // eps_ab = 2 pi^2 (f'_b v_a + f'_a v_b) H ; K = sum_{a<=b<Dim} fac sigma*_ab eps_ab / H
    double K = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int bb = a; bb < 3; ++bb) {
            if (bb >= ep.dim) continue;
            const double fac = (a != bb) ? 2.0 : 1.0;
            K += fac * ep.sig[a * 3 + bb] * 2.0 * kPi * kPi * (fm[bb] * v[a] + fm[a] * v[bb]);
        }
    return K;
}

// K(k) depends only on the frequencies and the material, so it is tabulated once (in FP64,
// stored in T)
template <typename T>
__global__ void k_kelastic_table(KElasticArgs<T> a, T* __restrict__ ktab) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        int kx, ky, kz;
        kcoords(a.g, i, a.nmain, kx, ky, kz);
        double f[3];
        ref_freq(a.g, kx, ky, kz, f);
        ktab[i] = (T)elastic_kernel(a.ep, f);
    }
}

template <typename T>
__global__ void k_kelastic(KElasticArgs<T> a) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        const double K = (double)a.ktab[i];
        const cpx<T> s = a.src[i];
        a.dst[i] = mk<T>((T)(K * (double)s.x), (T)(K * (double)s.y));
    }
}

template <typename T>
__global__ void k_sqsum(const T* w0, const T* w1, const T* w2, int nw, T factor, T* out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        T acc = w0[i] * w0[i];
        if (nw > 1) acc += w1[i] * w1[i];
        if (nw > 2) acc += w2[i] * w2[i];
        out[i] = factor * acc;
    }
}

template <typename T>
__global__ void k_copy(const T* src, T* dst, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// each thread moves 4 float4 per iteration with all loads issued before the stores
__global__ void __launch_bounds__(256) k_copy4(const float4* __restrict__ src, float4* __restrict__ dst, long n4) {
    const long stride = (long)gridDim.x * blockDim.x;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const float4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n4; i += stride) dst[i] = src[i];
}



constexpr int kRedThreads = 256;

template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_real_sums(const T* x, long n, double offset, int npoly,
                                                            const double* poly, double* pf, double* px) {
    __shared__ double red[kRedThreads];
    KSum sf, sx;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double c = (double)x[i] + offset;
        double f = 0.0;
        for (int q = npoly - 1; q >= 0; --q) f = f * c + poly[q];
        sf.add(f);
        sx.add(c);
    }
    const double a = block_sum(sf.get(), red);
    const double b = block_sum(sx.get(), red);
    if (threadIdx.x == 0) {
        pf[blockIdx.x] = a;
        px[blockIdx.x] = b;
    }
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_kgrad_energy(const cpx<T>* C, KGrid g, long nmain, long nk,
                                                               double* partial) {
    __shared__ double red[kRedThreads];
    KSum s;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nk; i += (long)gridDim.x * blockDim.x) {
        int kx, ky, kz;
        kcoords(g, i, nmain, kx, ky, kz);
        double f[3];
        ref_freq(g, kx, ky, kz, f);
        const double f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
        const double w = (kx == 0 || 2 * kx == g.nx) ? 1.0 : 2.0;
        const cpx<T> c = C[i];
        s.add(w * 4.0 * kPi * kPi * f2 * ((double)c.x * (double)c.x + (double)c.y * (double)c.y));
    }
    const double a = block_sum(s.get(), red);
    if (threadIdx.x == 0) partial[blockIdx.x] = a;
}

__global__ void __launch_bounds__(kRedThreads) k_reduce2(const double* a, long na, double ca, const double* b,
                                                          long nb, double cb, double* out) {
    __shared__ double red[kRedThreads];
    KSum sa, sb;
    for (long i = threadIdx.x; i < na; i += blockDim.x) sa.add(a[i]);
    for (long i = threadIdx.x; i < nb; i += blockDim.x) sb.add(b[i]);
    const double x = block_sum(sa.get(), red);
    const double y = block_sum(sb.get(), red);
    if (threadIdx.x == 0) out[0] = ca * x + cb * y;
}

__device__ __forceinline__ void adj_coeffs(const KGrid& g, long i, long nmain, double dt, const KPoly& den,
                                           const KPoly& lin, int src_lap, const double* filter, double& L, double& A,
                                           double& Ainv, double& m1, double& m2, double& w) {
    int kx, ky, kz;
    kcoords(g, i, nmain, kx, ky, kz);
    double f[3];
    ref_freq(g, kx, ky, kz, f);
    const double f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    L = -(4.0 * kPi * kPi) * f2;
    double D = 0.0, E = 0.0;
    for (int t = 0; t < den.n; ++t) D += den.coef[t] * ipow(L, den.pow[t]);
    for (int t = 0; t < lin.n; ++t) E += lin.coef[t] * ipow(L, lin.pow[t]);
    Ainv = 1.0 / (1.0 - dt * D);
    A = Ainv;
    if (filter) A *= vandeven(filter, sqrt(f2) * 2.0 / kPi);
    m1 = A * (1.0 + dt * E);
    m2 = A * dt * ipow(L, src_lap);
    w = (kx == 0 || 2 * kx == g.nx) ? 1.0 : 2.0;
}

template <typename T>
__global__ void k_adj_mult(AdjArgs<T> a) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        double L, A, Ainv, m1, m2, w;
        adj_coeffs(a.g, i, a.nmain, a.dt, a.denom, a.lin, a.src_lap, a.filter, L, A, Ainv, m1, m2, w);
        const cpx<T> l = a.lam[i];
        a.P[i] = mk<T>((T)(m1 * (double)l.x), (T)(m1 * (double)l.y));
        a.Q[i] = mk<T>((T)(m2 * (double)l.x), (T)(m2 * (double)l.y));
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_adj_grad(AdjArgs<T> a) {
    __shared__ double red[256];
    double acc[kMaxAdjScalars];
    for (int s = 0; s < kMaxAdjScalars; ++s) acc[s] = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        double L, A, Ainv, m1, m2, w;
        adj_coeffs(a.g, i, a.nmain, a.dt, a.denom, a.lin, a.src_lap, a.filter, L, A, Ainv, m1, m2, w);
        const cpx<T> l = a.lam[i];
        a.P[i] = mk<T>((T)(m1 * (double)l.x), (T)(m1 * (double)l.y));
        a.Q[i] = mk<T>((T)(m2 * (double)l.x), (T)(m2 * (double)l.y));
        double c1x = 0.0, c1y = 0.0, c0x = 0.0, c0y = 0.0;
        if (a.C1) {
            c1x = a.cscale * (double)a.C1[i].x;
            c1y = a.cscale * (double)a.C1[i].y;
            if (i == 0 && a.mean1) c1x += a.cscale * *a.mean1;
        }
        if (a.C0) {
            c0x = a.cscale * (double)a.C0[i].x;
            c0y = a.cscale * (double)a.C0[i].y;
            if (i == 0 && a.mean0) c0x += a.cscale * *a.mean0;
        }
        // Re(conj(l) z) = l.x z.x + l.y z.y
        const double lc1 = (double)l.x * c1x + (double)l.y * c1y;
        const double lc0 = (double)l.x * c0x + (double)l.y * c0y;
#pragma unroll
        for (int s = 0; s < kMaxAdjScalars; ++s) {
            if (s >= a.nscal) break;
            const double* dc = a.dcoef + (long)s * 2 * kMaxKTerms;
            double dD = 0.0, dE = 0.0;
            for (int t = 0; t < a.denom.n; ++t) dD += dc[t] * ipow(L, a.denom.pow[t]);
            for (int t = 0; t < a.lin.n; ++t) dE += dc[kMaxKTerms + t] * ipow(L, a.lin.pow[t]);
            acc[s] += w * (Ainv * a.dt * dD * lc1 + A * a.dt * dE * lc0);
        }
    }
#pragma unroll
    for (int s = 0; s < kMaxAdjScalars; ++s) {
        if (s >= a.nscal) break;  // uniform over the block
        const double tot = block_sum(acc[s], red);
        if (threadIdx.x == 0) a.partial[(long)s * gridDim.x + blockIdx.x] += tot * a.inv_n;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_adj_lambda(const T* p, const T* q, const T* c, int npoly, const double* dpoly,
                                                    T* out, long n, int nscal, const double* spoly, double* partial) {
    __shared__ double red[256];
    double acc[kMaxAdjScalars];
    for (int s = 0; s < kMaxAdjScalars; ++s) acc[s] = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double x = (double)c[i];
        double d = 0.0;
        for (int k = npoly - 1; k >= 0; --k) d = d * x + dpoly[k];
        const double qv = q ? (double)q[i] : 1.0;
        out[i] = (T)((p ? (double)p[i] : 0.0) + d * qv);
#pragma unroll
        for (int s = 0; s < kMaxAdjScalars; ++s) {
            if (s >= nscal) break;
            const double* sp = spoly + (long)s * kMaxPoly;
            double h = 0.0;
            for (int k = kMaxPoly - 1; k >= 0; --k) h = h * x + sp[k];
            acc[s] += qv * h;
        }
    }
#pragma unroll
    for (int s = 0; s < kMaxAdjScalars; ++s) {
        if (s >= nscal) break;  // uniform over the block
        const double tot = block_sum(acc[s], red);
        if (threadIdx.x == 0) partial[(long)s * gridDim.x + blockIdx.x] += tot;
    }
}

__global__ void __launch_bounds__(256) k_accumulate(const double* partial, int nscal, long nblocks, double* acc) {
    __shared__ double red[256];
    for (int s = 0; s < nscal; ++s) {
        KSum k;
        for (long b = threadIdx.x; b < nblocks; b += blockDim.x) k.add(partial[(long)s * nblocks + b]);
        const double tot = block_sum(k.get(), red);
        if (threadIdx.x == 0) acc[s] += tot;
    }
}

inline unsigned grid_for(long n, int thr = 256) {
    long g = (n + thr - 1) / thr;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// ---------------------------------------------------------------------------
// SDD (sdd.go:159-301)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sdd_weight(const KGrid& g, long i, long nmain) {
    int kx, ky, kz;
    kcoords(g, i, nmain, kx, ky, kz);
    return (kx == 0 || 2 * kx == g.nx) ? 1.0 : 2.0;
}

__device__ __forceinline__ double sdd_D(const SddArgs& a, long i) {
    int kx, ky, kz;
    kcoords(a.g, i, a.nmain, kx, ky, kz);
    double f[3];
    ref_freq(a.g, kx, ky, kz, f);
    const double f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    return denominator(a.denom, a.sv, a.hs, a.g.ndim, f, -(4.0 * kPi * kPi) * f2, f2);
}

__device__ __forceinline__ double block_reduce(double v, bool is_max) {
    __shared__ double red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = is_max ? fmax(red[threadIdx.x], red[threadIdx.x + s])
                                                            : red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(256) k_sdd_dot(SddArgs a) {
    double acc = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        const cpx<double> r0 = a.rs0[i], r1 = a.re0[i], o = a.ov[i];
        const double wx = a.alpha * r0.x + (1.0 - a.alpha) * r1.x, wy = a.alpha * r0.y + (1.0 - a.alpha) * r1.y;
        acc += sdd_weight(a.g, i, a.nmain) * (wx * o.x + wy * o.y);
    }
    const double t = block_reduce(acc, false);
    if (threadIdx.x == 0) a.part[blockIdx.x] = t;
}

// d = k + dt (work - 2 ov dot): the Euler update after the Householder reflection (sdd.go:218-225)
__device__ __forceinline__ cpx<double> sdd_d(const SddArgs& a, long i, double dot) {
    const cpx<double> r0 = a.rs0[i], r1 = a.re0[i], o = a.ov[i], k = a.k[i];
    const double wx = a.alpha * r0.x + (1.0 - a.alpha) * r1.x - 2.0 * o.x * dot;
    const double wy = a.alpha * r0.y + (1.0 - a.alpha) * r1.y - 2.0 * o.y * dot;
    return mk<double>(k.x + a.dt * wx, k.y + a.dt * wy);
}

__global__ void __launch_bounds__(256) k_sdd_sm(SddArgs a) {
    const double dot = a.sc[0];
    double accd = 0.0, accs = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        const double D = sdd_D(a, i);
        const double inv = 1.0 / (1.0 - a.dt * D);
        const cpx<double> o = a.ov[i], d = sdd_d(a, i, dot);
        const double w = sdd_weight(a.g, i, a.nmain) * 2.0 * D * a.dt * a.inv_n * inv;
        accd += w * (o.x * o.x + o.y * o.y);
        accs += w * (o.x * d.x + o.y * d.y);
    }
    const double td = block_reduce(accd, false);
    const double ts = block_reduce(accs, false);
    if (threadIdx.x == 0) {
        a.part[blockIdx.x] = td;
        a.part[kSddBlocks + blockIdx.x] = ts;
    }
}

__global__ void __launch_bounds__(256) k_sdd_update(SddArgs a) {
    const double dot = a.sc[0], den = a.sc[1], proj = a.sc[2];
    double accp = 0.0, accm = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        const double D = sdd_D(a, i);
        const double inv = 1.0 / (1.0 - a.dt * D);
        const cpx<double> o = a.ov[i], d = sdd_d(a, i, dot), k0 = a.k[i];
        const double nx = inv * d.x - inv * o.x * proj / den, ny = inv * d.y - inv * o.y * proj / den;
        const double dx = (nx - k0.x) / a.dt, dy = (ny - k0.y) / a.dt;
        const double q = dx * dx + dy * dy;
        accp += sdd_weight(a.g, i, a.nmain) * q * a.inv_n;
        accm = fmax(accm, sqrt(q));
        a.k[i] = mk<double>(nx, ny);
    }
    const double tp = block_reduce(accp, false);
    const double tm = block_reduce(accm, true);
    if (threadIdx.x == 0) {
        a.part[blockIdx.x] = tp;
        a.part[kSddBlocks + blockIdx.x] = tm;
    }
}

__global__ void __launch_bounds__(256) k_sdd_torque(SddArgs a, int first) {
    double acc = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        const double D = sdd_D(a, i);
        const cpx<double> r = a.torque[i], e = a.re[i], o = a.ov[i];
        const double tx = r.x - e.x - D * o.x * a.len, ty = r.y - e.y - D * o.y * a.len;
        a.torque[i] = mk<double>(tx, ty);
        acc += sdd_weight(a.g, i, a.nmain) * (tx * o.x + ty * o.y);
    }
    const double t = block_reduce(acc, false);
    if (threadIdx.x == 0) a.part[blockIdx.x] = first ? t : a.part[blockIdx.x] + t;
}

__global__ void __launch_bounds__(256) k_sdd_project(SddArgs a) {
    const double p = a.sc[3];
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.nk; i += (long)gridDim.x * blockDim.x) {
        const cpx<double> t = a.torque[i], o = a.ov[i];
        a.torque[i] = mk<double>(t.x - o.x * p, t.y - o.y * p);
    }
}

__global__ void __launch_bounds__(256) k_sdd_orient(double* v, const double* tau, long n, double c, double* part) {
    double s2 = 0.0, mx = 0.0, sd = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double t = tau[i], v0 = v[i];
        sd += t * v0;
        mx = fmax(mx, fabs(t));
        const double v1 = v0 - c * t;
        v[i] = v1;
        s2 += v1 * v1;
    }
    const double a = block_reduce(s2, false);
    const double b = block_reduce(mx, true);
    const double d = block_reduce(sd, false);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[kSddBlocks + blockIdx.x] = b;
        part[2 * kSddBlocks + blockIdx.x] = d;
    }
}

// mode 0: sc[0] = sum part / N (a0 = 1/N)
// mode 1: sc[1] = 1 + sum part, sc[2] = sum part[kSddBlocks..]
// mode 2: mon[1] += sum part, mon[0] = max(mon[0], max part[kSddBlocks..])   (mode 5 first resets them)
// mode 3: sc[3] = sum part / N (a0 = 1/N)
// mode 4: mon[2] = max part[kSddBlocks..], mon[5] = sum part[2 kSddBlocks..], sc[4] = sum part (|v|^2);
//         mon[1] = sqrt(mon[1] / a1) (a1 = nf N)
// mode 6: field norm: fn = sum of the a0 values part[0..a0); mon[4] = mon[3] - fn; mon[3] = fn
__global__ void k_sdd_scalars(int mode, const double* part, double* sc, double* mon, double a0, double a1) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s0 = 0.0, s1 = 0.0, m1 = 0.0, s2 = 0.0;
    const int nparts = (mode == 6) ? (int)a0 : kSddBlocks;
    for (int b = 0; b < nparts; ++b) {
        s0 += part[b];
        if (mode == 1) s1 += part[kSddBlocks + b];
        if (mode == 2 || mode == 5 || mode == 4) m1 = fmax(m1, part[kSddBlocks + b]);
        if (mode == 4) s2 += part[2 * kSddBlocks + b];
    }
    if (mode == 0) sc[0] = s0 * a0;
    if (mode == 1) {
        sc[1] = 1.0 + s0;
        sc[2] = s1;
    }
    if (mode == 5) {
        mon[1] = 0.0;
        mon[0] = 0.0;
    }
    if (mode == 2 || mode == 5) {
        mon[1] += s0;
        mon[0] = fmax(mon[0], m1);
    }
    if (mode == 3) sc[3] = s0 * a0;
    if (mode == 4) {
        mon[2] = m1;
        mon[5] = s2;
        sc[4] = s0;
        mon[1] = sqrt(mon[1] / a1);
    }
    if (mode == 6) {
        mon[4] = mon[3] - s0;
        mon[3] = s0;
    }
}

}  // namespace

hipError_t launch_sdd_dot(const SddArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_sdd_dot, dim3(kSddBlocks), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_sdd_sm(const SddArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_sdd_sm, dim3(kSddBlocks), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_sdd_update(const SddArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_sdd_update, dim3(kSddBlocks), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_sdd_torque(const SddArgs& a, bool first, hipStream_t s) {
    hipLaunchKernelGGL(k_sdd_torque, dim3(kSddBlocks), dim3(256), 0, s, a, first ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch_sdd_project(const SddArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_sdd_project, dim3(kSddBlocks), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_sdd_orient(double* v, const double* tau, long n, double c, double* part, hipStream_t s) {
    hipLaunchKernelGGL(k_sdd_orient, dim3(kSddBlocks), dim3(256), 0, s, v, tau, n, c, part);
    return hipGetLastError();
}
hipError_t launch_sdd_scalars(int mode, const double* part, double* sc, double* mon, double a0, double a1,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_sdd_scalars, dim3(1), dim3(64), 0, s, mode, part, sc, mon, a0, a1);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_derived(const DerivedArgs<T>& a, hipStream_t s) {
    hipLaunchKernelGGL(k_derived<T>, dim3(grid_for(a.n)), dim3(256), 0, s, a);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_kupdate(const KUpdateArgs<T>& a, hipStream_t s) {
    hipLaunchKernelGGL(k_kupdate<T>, dim3(grid_for(a.nk)), dim3(256), 0, s, a);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_rk4_stage(const Rk4StageArgs<T>& a, hipStream_t s) {
    constexpr int EPT = kStageEpt;
    const long blocks = (a.nk + kStageThreads * EPT - 1) / (kStageThreads * EPT);
    if (a.stab) {
        for (int t = 0; t < a.nterm; ++t)
            if (a.tmode[t] != 0) return hipErrorInvalidValue;  // needs the frequency vector
        // FP64 operands take twice the registers: two elements per thread keep 4 waves per SIMD
        constexpr int ET = sizeof(T) == 8 ? 2 : 4;
        const long bt = (a.nk + kStageThreads * ET - 1) / (kStageThreads * ET);
        if (a.need_l)
            hipLaunchKernelGGL((k_rk4_stage_tab<T, ET, true>), dim3((unsigned)bt), dim3(kStageThreads), 0, s, a);
        else
            hipLaunchKernelGGL((k_rk4_stage_tab<T, ET, false>), dim3((unsigned)bt), dim3(kStageThreads), 0, s, a);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_rk4_stage<T, EPT>), dim3((unsigned)blocks), dim3(kStageThreads), 0, s, a);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_stage_denoms(const StageDenomArgs<T>& a, hipStream_t s) {
    if (!a.out) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_stage_denoms<T>, dim3(grid_for(a.nk)), dim3(256), 0, s, a);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_etd_residual(const EtdArgs<T>& a, hipStream_t s) {
    hipLaunchKernelGGL(k_etd_residual<T>, dim3(grid_for(a.nk)), dim3(256), 0, s, a);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_volcons_update(const double* px, long nb, const T* ind, double dt, double* st, hipStream_t s) {
    hipLaunchKernelGGL(k_volcons_update<T>, dim3(1), dim3(64), 0, s, px, nb, ind, dt, st);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_kgrad(const KGradArgs<T>& a, int ndim, hipStream_t s) {
    hipLaunchKernelGGL(k_kgrad<T>, dim3(grid_for(a.nk)), dim3(256), 0, s, a, ndim);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_kelastic_table(const KElasticArgs<T>& a, T* ktab, hipStream_t s) {
    hipLaunchKernelGGL(k_kelastic_table<T>, dim3(grid_for(a.nk)), dim3(256), 0, s, a, ktab);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_kelastic(const KElasticArgs<T>& a, hipStream_t s) {
    if (!a.ktab) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_kelastic<T>, dim3(grid_for(a.nk)), dim3(256), 0, s, a);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_sqsum(const T* const* w, int nw, T factor, T* out, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_sqsum<T>, dim3(grid_for(n)), dim3(256), 0, s, w[0], nw > 1 ? w[1] : w[0],
                       nw > 2 ? w[2] : w[0], nw, factor, out, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_copy(const T* src, T* dst, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_copy<T>, dim3(grid_for(n)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_real_sums(const T* x, long n, double offset, int npoly, const double* poly, double* pf, double* px,
                            long nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_real_sums<T>, dim3((unsigned)nblocks), dim3(kRedThreads), 0, s, x, n, offset, npoly, poly,
                       pf, px);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_kgrad_energy(const cpx<T>* C, KGrid g, long nmain, long nk, double* partial, long nblocks,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_kgrad_energy<T>, dim3((unsigned)nblocks), dim3(kRedThreads), 0, s, C, g, nmain, nk, partial);
    return hipGetLastError();
}
// achievable-bandwidth probes: 16-byte accesses, four per thread in flight, grid-stride over
// 16384 blocks; the copy streams with nontemporal loads and stores (measured here as the
// fastest copy form: 6.1 TB/s against 5.7 TB/s with ordinary stores on a 1 GiB array)
typedef float gpf_f4v __attribute__((ext_vector_type(4)));
constexpr int kBwU = 4;

__global__ void __launch_bounds__(256) k_bw_copy(const gpf_f4v* __restrict__ s, gpf_f4v* __restrict__ d, long n4) {
    const long stride = (long)gridDim.x * blockDim.x * kBwU;
    for (long i = (long)blockIdx.x * blockDim.x * kBwU + threadIdx.x; i + (kBwU - 1) * (long)blockDim.x < n4;
         i += stride) {
        gpf_f4v v[kBwU];
#pragma unroll
        for (int u = 0; u < kBwU; ++u) v[u] = __builtin_nontemporal_load(&s[i + u * blockDim.x]);
#pragma unroll
        for (int u = 0; u < kBwU; ++u) __builtin_nontemporal_store(v[u], &d[i + u * blockDim.x]);
    }
}

__global__ void __launch_bounds__(256) k_bw_read(const gpf_f4v* __restrict__ s, float* __restrict__ sink, long n4) {
    const long stride = (long)gridDim.x * blockDim.x * kBwU;
    float acc = 0.f;
    for (long i = (long)blockIdx.x * blockDim.x * kBwU + threadIdx.x; i + (kBwU - 1) * (long)blockDim.x < n4;
         i += stride) {
        gpf_f4v v[kBwU];
#pragma unroll
        for (int u = 0; u < kBwU; ++u) v[u] = s[i + u * blockDim.x];
#pragma unroll
        for (int u = 0; u < kBwU; ++u) acc += v[u].x + v[u].y + v[u].z + v[u].w;
    }
    if (acc == -1.0f) sink[0] = acc;  // keeps the loads; the arrays hold zeros
}

__global__ void __launch_bounds__(256) k_bw_write(gpf_f4v* __restrict__ d, long n4) {
    const long stride = (long)gridDim.x * blockDim.x * kBwU;
    const gpf_f4v z = {0.f, 0.f, 0.f, 0.f};
    for (long i = (long)blockIdx.x * blockDim.x * kBwU + threadIdx.x; i + (kBwU - 1) * (long)blockDim.x < n4;
         i += stride) {
#pragma unroll
        for (int u = 0; u < kBwU; ++u) d[i + u * blockDim.x] = z;
    }
}

// kind 0 copy src -> dst, 1 read src, 2 write dst; n floats, a multiple of 4 * 256 * kBwU
hipError_t launch_bw(int kind, const float* src, float* dst, long n, hipStream_t s) {
    const long n4 = n / 4;
    const dim3 grid(16384), block(256);
    if (kind == 0)
        hipLaunchKernelGGL(k_bw_copy, grid, block, 0, s, reinterpret_cast<const gpf_f4v*>(src),
                           reinterpret_cast<gpf_f4v*>(dst), n4);
    else if (kind == 1)
        hipLaunchKernelGGL(k_bw_read, grid, block, 0, s, reinterpret_cast<const gpf_f4v*>(src), dst, n4);
    else
        hipLaunchKernelGGL(k_bw_write, grid, block, 0, s, reinterpret_cast<gpf_f4v*>(dst), n4);
    return hipGetLastError();
}

hipError_t launch_copy_bw(const float* src, float* dst, long n, hipStream_t s) {
    const long n4 = n / 4;
    long g = (n4 + 1023) / 1024;
    if (g > 16384) g = 16384;
    hipLaunchKernelGGL(k_copy4, dim3((unsigned)g), dim3(256), 0, s, reinterpret_cast<const float4*>(src),
                       reinterpret_cast<float4*>(dst), n4);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_adj_mult(const AdjArgs<T>& a, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_mult<T>, dim3(grid_for(a.nk)), dim3(256), 0, s, a);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_adj_grad(const AdjArgs<T>& a, long nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_grad<T>, dim3((unsigned)nblocks), dim3(256), 0, s, a);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_adj_lambda(const T* p, const T* q, const T* c, int npoly, const double* dpoly, T* out, long n,
                             int nscal, const double* spoly, double* partial, long nblocks, hipStream_t s) {
    const unsigned grid = nscal > 0 ? (unsigned)nblocks : grid_for(n);
    hipLaunchKernelGGL(k_adj_lambda<T>, dim3(grid), dim3(256), 0, s, p, q, c, npoly, dpoly, out, n, nscal, spoly,
                       partial);
    return hipGetLastError();
}
hipError_t launch_accumulate(const double* partial, int nscal, long nblocks, double* acc, hipStream_t s) {
    hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(256), 0, s, partial, nscal, nblocks, acc);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kRedThreads) k_series_append(const double* bulk, long nbulk, const double* eg,
                                                                 long neg, double ceg, const double* mass,
                                                                 double* series, long* counter) {
    series_slot(bulk, mass, nbulk, eg, neg, ceg, series, counter);
}

hipError_t launch_series_append(const double* bulk, long nbulk, const double* eg, long neg, double ceg,
                                const double* mass, double* series, long* counter, hipStream_t s) {
    hipLaunchKernelGGL(k_series_append, dim3(1), dim3(kRedThreads), 0, s, bulk, nbulk, eg, neg, ceg, mass, series,
                       counter);
    return hipGetLastError();
}

hipError_t launch_reduce2(const double* a, long na, double ca, const double* b, long nb, double cb, double* out,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kRedThreads), 0, s, a, na, ca, b, nb, cb, out);
    return hipGetLastError();
}

#define GPF_INST(T)                                                                                            \
    template hipError_t launch_derived<T>(const DerivedArgs<T>&, hipStream_t);                                \
    template hipError_t launch_kupdate<T>(const KUpdateArgs<T>&, hipStream_t);                                \
    template hipError_t launch_etd_residual<T>(const EtdArgs<T>&, hipStream_t);                               \
    template hipError_t launch_rk4_stage<T>(const Rk4StageArgs<T>&, hipStream_t);                             \
    template hipError_t launch_stage_denoms<T>(const StageDenomArgs<T>&, hipStream_t);                        \
    template hipError_t launch_volcons_update<T>(const double*, long, const T*, double, double*, hipStream_t); \
    template hipError_t launch_kgrad<T>(const KGradArgs<T>&, int, hipStream_t);                               \
    template hipError_t launch_kelastic<T>(const KElasticArgs<T>&, hipStream_t);                              \
    template hipError_t launch_kelastic_table<T>(const KElasticArgs<T>&, T*, hipStream_t);                    \
    template hipError_t launch_sqsum<T>(const T* const*, int, T, T*, long, hipStream_t);                      \
    template hipError_t launch_copy<T>(const T*, T*, long, hipStream_t);                                      \
    template hipError_t launch_real_sums<T>(const T*, long, double, int, const double*, double*, double*,     \
                                            long, hipStream_t);                                               \
    template hipError_t launch_kgrad_energy<T>(const cpx<T>*, KGrid, long, long, double*, long, hipStream_t); \
    template hipError_t launch_adj_mult<T>(const AdjArgs<T>&, hipStream_t);                                  \
    template hipError_t launch_adj_grad<T>(const AdjArgs<T>&, long, hipStream_t);                            \
    template hipError_t launch_adj_lambda<T>(const T*, const T*, const T*, int, const double*, T*, long, int, \
                                             const double*, double*, long, hipStream_t);
GPF_INST(float)
GPF_INST(double)
#undef GPF_INST

}  // namespace gpf
