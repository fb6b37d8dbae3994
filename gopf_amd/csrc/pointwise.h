// pointwise.h -- pointwise / k-space kernels of the general (multi-field) stepper.
#pragma once

#include <hip/hip_runtime.h>

#include "spectral.h"

namespace gpf {

constexpr int kMaxFields = 8;
constexpr int kMaxDTerms = 16;
constexpr int kMaxUTerms = 10;

// real-space derived field: out = sum_t coef_t prod_f x_f^pow[t][f]
template <typename T>
struct DerivedArgs {
    const T* f[kMaxFields] = {};
    T* out = nullptr;
    int nf = 0;
    int nterm = 0;
    double coef[kMaxDTerms];
    double pw[kMaxDTerms][kMaxFields];
    long n = 0;
    // dense form for at most two fields and integer powers 0..3:
    // out = sum_ij dc[i][j] f0^i f1^j (f1 = 0 when nf == 1)
    int dense = 0;
    double dc[4][4] = {};
    // Gaussian noise (WhiteNoise.Generate, noise.go:19-22, and ConservativeNoise's current
    // fields, noise.go:86-100): out[i] = noise_std * z, z standard normal from a counter-based
    // generator keyed by seed on (i, counter, stream); counter advances at every sync
    int noise = 0;
    double noise_std = 0.0;
    unsigned long long seed = 0;
    unsigned stream = 0, counter = 0;
};

// Philox4x32-10 (Salmon et al., SC'11) and the normal draw of one node, shared by the
// device kernels and the host tests
struct NoiseDraw {
    __host__ __device__ static inline unsigned mulhi(unsigned a, unsigned b) {
        return (unsigned)(((unsigned long long)a * b) >> 32);
    }
    __host__ __device__ static inline void philox(unsigned c[4], unsigned k0, unsigned k1) {
        for (int r = 0; r < 10; ++r) {
            if (r > 0) {
                k0 += 0x9E3779B9u;
                k1 += 0xBB67AE85u;
            }
            const unsigned hi0 = mulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
            const unsigned hi1 = mulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
            const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
            c[0] = n0;
            c[1] = lo1;
            c[2] = n2;
            c[3] = lo0;
        }
    }
    // two 53-bit uniforms in (0, 1] from the four words, then Box-Muller
    __host__ __device__ static inline double normal(long i, unsigned counter, unsigned stream,
                                                    unsigned long long seed) {
        unsigned c[4] = {(unsigned)((unsigned long long)i & 0xffffffffu), (unsigned)((unsigned long long)i >> 32),
                         counter, stream};
        philox(c, (unsigned)(seed & 0xffffffffu), (unsigned)(seed >> 32));
        const double s53 = 1.0 / 9007199254740992.0;
        const double u1 = ((double)((((unsigned long long)(c[0] >> 5)) << 26) | (c[1] >> 6)) + 0.5) * s53;
        const double u2 = ((double)((((unsigned long long)(c[2] >> 5)) << 26) | (c[3] >> 6)) + 0.5) * s53;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
    }
};

// fill DerivedArgs::dc when the term list has the dense form; returns a.dense
template <typename T>
inline int make_dense_derived(DerivedArgs<T>& a) {
    a.dense = 0;
    if (a.nf < 1 || a.nf > 2) return 0;
    double dc[4][4] = {};
    for (int t = 0; t < a.nterm; ++t) {
        int e[2] = {0, 0};
        for (int f = 0; f < a.nf; ++f) {
            const double p = a.pw[t][f];
            if (p < 0.0 || p > 3.0 || p != (double)(int)p) return 0;
            e[f] = (int)p;
        }
        dc[e[0]][e[1]] += a.coef[t];
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) a.dc[i][j] = dc[i][j];
    return a.dense = 1;
}

// implicit SpectralViscosity terms of a denominator (spectralViscosity.go:44-52 with the
// constructFunc prefixes): sum_t coef_t L^lap_t Q(|f|, thr_t) |f|^pw_t, |f| in cycles per
// sample; coef_t carries the prefix sign and -Eps
constexpr int kMaxSv = 4;
struct SvPoly {
    int n = 0;
    double coef[kMaxSv];
    double thr[kMaxSv];
    int pw[kMaxSv];
    int lap[kMaxSv];
};

// implicit TensorialHessian terms (tensorialHessian.go:37-60 with the constructFunc
// prefixes): sum_t coef_t L^lap_t [sum_j -4 pi^2 f_j^2 K_jj + sum_{j<k} -8 pi^2 f_j f_k K_jk]
constexpr int kMaxHess = 2;
struct HessPoly {
    int n = 0;
    double coef[kMaxHess];
    int lap[kMaxHess];
    double K[kMaxHess][9];  // row-major, leading dimension kd
    int kd[kMaxHess];
};

// half-spectrum k-space update
//   dst = filt * (alpha * cur + beta * sum_t coef_t L^pow_t B_t) / (1 - gamma * D(k))
// B_t == null means the brick is the constant 1.
template <typename T>
struct KUpdateArgs {
    cpx<T>* dst = nullptr;
    const cpx<T>* cur = nullptr;       // may be null (alpha ignored)
    int nterm = 0;
    const cpx<T>* brick[kMaxUTerms] = {};
    double coef[kMaxUTerms];
    int pw[kMaxUTerms];
    const double* dmul[kMaxUTerms] = {};  // optional device-resident factor of a term (null = 1)
    // tmode 1: the multiplier is i 2 sin(pi f_comp), 0 on Nyquist rows (ConservativeNoise's
    // discrete divergence, noise.go:60-78), comp = tcomp
    int tmode[kMaxUTerms] = {};
    int tcomp[kMaxUTerms] = {};
    double alpha = 1.0, beta = 0.0, gamma = 0.0;
    KPoly denom;
    SvPoly sv;
    HessPoly hs;
    const double* filter = nullptr;
    KGrid g;
    long nmain = 0, nk = 0;            // main elements; total (main + nyq)
};

// residual of the exponential-integrator step of ImplicitEuler on a half spectrum:
//   dst = cur - (orig e^{D dt} + I(D, R, rprev)),  R = sum_t coef_t L^pow_t B_t
// with I the integral over the step of e^{D (dt - s)} times the explicit part interpolated
// linearly from rprev (s = 0) to R (s = dt); B_t == null means the brick is the constant 1
template <typename T>
struct EtdArgs {
    cpx<T>* dst = nullptr;
    const cpx<T>* cur = nullptr;
    const cpx<T>* orig = nullptr;
    const cpx<T>* rprev = nullptr;
    int nterm = 0;
    const cpx<T>* brick[kMaxUTerms] = {};
    double coef[kMaxUTerms];
    int pw[kMaxUTerms];
    const double* dmul[kMaxUTerms] = {};
    int tmode[kMaxUTerms] = {};
    int tcomp[kMaxUTerms] = {};
    KPoly denom;
    SvPoly sv;
    HessPoly hs;
    double dt = 0.0;
    KGrid g;
    long nmain = 0, nk = 0;
};

// one RK4 stage of RK4.Step (rk4.go:29-127) per field, with the stage's right-hand side
// R = sum_t coef_t L^pow_t B_t evaluated in the kernel instead of stored:
//   stage 0:    init = k; final = k + fp dt R;    kn = (k + fc dt R) / (1 - fc dt D)
//   stage 1, 2: final += fp dt R;                 kn = (init + fc dt R) / (1 - fc dt D)
//   stage 3:    kn = filt (final + fp dt R) / (1 - dt D)
// (fp, fc: the stage's PrepareNextCorrection and correction factors); kn is a separate
// array so that later fields of the stage still read this field's brick unchanged
template <typename T>
struct Rk4StageArgs {
    const cpx<T>* k = nullptr;
    cpx<T>* kn = nullptr;
    cpx<T>* init = nullptr;
    cpx<T>* final = nullptr;
    int stage = 0;
    double dt = 0.0, fp = 0.0, fc = 0.0;
    int nterm = 0;
    const cpx<T>* brick[kMaxUTerms] = {};
    double coef[kMaxUTerms];
    int pw[kMaxUTerms];
    const double* dmul[kMaxUTerms] = {};
    int tmode[kMaxUTerms] = {};
    int tcomp[kMaxUTerms] = {};
    KPoly denom;
    SvPoly sv;
    HessPoly hs;
    const double* filter = nullptr;
    KGrid g;
    long nmain = 0, nk = 0;
    // optional table of this stage's factor s(k) = filt / (1 - gamma dt D(k)) (built once per
    // plan by launch_stage_denoms, in FP64 and stored in T); the kernel then computes no
    // denominator and, when no term has a LAP power (need_l = 0), no frequency at all
    const T* stab = nullptr;
    int need_l = 1;
};
template <typename T> hipError_t launch_rk4_stage(const Rk4StageArgs<T>& a, hipStream_t s);

// the factor table of one RK4 stage kind: out[k] = filt(k) / (1 - gamma D(k)) with the
// TensorialHessian rows averaged over k and -k, as k_rk4_stage forms it (filter optional)
template <typename T>
struct StageDenomArgs {
    KPoly denom;
    SvPoly sv;
    HessPoly hs;
    double gamma = 0.0;
    const double* filter = nullptr;
    KGrid g;
    long nmain = 0, nk = 0;
    T* out = nullptr;
};
template <typename T> hipError_t launch_stage_denoms(const StageDenomArgs<T>& a, hipStream_t s);

// spectral gradient components: out_d = i 2 pi f_d' src   (f_d' = 0 at f_d = +0.5)
template <typename T>
struct KGradArgs {
    const cpx<T>* src = nullptr;
    cpx<T>* out[3] = {};
    KGrid g;
    long nmain = 0, nk = 0;
};

// Khachaturyan kernel: dst = K(k) src with K the scalar elastic kernel
struct ElasticParams {
    double C[81];
    double sig[9];   // effective stress C : misfit
    int dim = 3;
};
template <typename T>
struct KElasticArgs {
    const cpx<T>* src = nullptr;
    cpx<T>* dst = nullptr;
    ElasticParams ep;
    KGrid g;
    long nmain = 0, nk = 0;
    const T* ktab = nullptr;       // K(k), filled once by launch_kelastic_table
};

// adjoint of the fused semi-implicit step (single field)
constexpr int kMaxAdjScalars = 8;
template <typename T>
struct AdjArgs {
    KGrid g;
    long nmain = 0, nk = 0;
    double dt = 0.0;
    KPoly denom, lin;          // D(k), E(k)
    int src_lap = 0;
    const double* filter = nullptr;
    // k_adj_mult: P = m1 Lam, Q = m2 Lam with m1 = A (1 + dt E), m2 = A dt L^src_lap, A = sigma / (1 - dt D)
    const cpx<T>* lam = nullptr;
    cpx<T>* P = nullptr;
    cpx<T>* Q = nullptr;
    // k_adj_grad: P and Q as k_adj_mult, and per scalar s the block sums of
    // w Re(conj(Lam) [dt dD_s C1 / (1 - dt D) + A dt dE_s C0]) / N added to partial[s][block], with
    // dD_s = sum_t dcoef[s][0][t] L^denom.pow[t], dE_s = sum_t dcoef[s][1][t] L^lin.pow[t] and the
    // transforms C1 = FFT(c_{n+1}), C0 = FFT(c_n) given as state spectra: C = cscale S, plus
    // N mean at k = 0 (mean1 / mean0 may be null)
    int nscal = 0;
    const double* dcoef = nullptr;  // device [nscal][2][kMaxKTerms]
    const cpx<T>* C1 = nullptr;
    const cpx<T>* C0 = nullptr;
    const double* mean1 = nullptr;
    const double* mean0 = nullptr;
    double cscale = 1.0;            // N
    double* partial = nullptr;      // [nscal][nblocks], accumulated across launches
    double inv_n = 1.0;
};

template <typename T> hipError_t launch_adj_mult(const AdjArgs<T>& a, hipStream_t s);
template <typename T> hipError_t launch_adj_grad(const AdjArgs<T>& a, long nblocks, hipStream_t s);
// out = p + h(c) q with h(c) = sum_k dpoly[k] c^k; a null p reads as 0 and a null q as 1
// (the adjoint's lam_n = p + g'(c_n) q, and plain polynomial evaluation).  With nscal > 0
// it also adds, per scalar s, the block sums of q(x) sum_k spoly[s][k] c(x)^k to
// partial[s][block] (spoly: device [nscal][kMaxPoly]); launched with nblocks blocks.
template <typename T> hipError_t launch_adj_lambda(const T* p, const T* q, const T* c, int npoly, const double* dpoly,
                                                   T* out, long n, int nscal, const double* spoly, double* partial,
                                                   long nblocks, hipStream_t s);
// acc[s] += sum_b partial[s][b]  (single block, deterministic order)
hipError_t launch_accumulate(const double* partial, int nscal, long nblocks, double* acc, hipStream_t s);

template <typename T> hipError_t launch_derived(const DerivedArgs<T>& a, hipStream_t s);
template <typename T> hipError_t launch_kupdate(const KUpdateArgs<T>& a, hipStream_t s);

// SDD.Step (sdd.go:159-301), one field's spectral part on the half spectrum (FP64).  Sums over
// the full spectrum of Hermitian products are Hermitian-weighted half sums.  Device scalars
// sc: [0] <work, ov> / N, [1] the Sherman-Morrison denominator, [2] its projection, [3] the
// torque projection; every kernel writes one partial per block to part (kSddBlocks blocks).
constexpr int kSddBlocks = 256;
struct SddArgs {
    KGrid g;
    long nmain = 0, nk = 0;
    double inv_n = 1.0;
    KPoly denom;          // D(k) of the field's equation
    SvPoly sv;
    HessPoly hs;
    double dt = 0.0, alpha = 0.5, len = 0.0;
    const cpx<double>* rs0 = nullptr;  // field 0's GetRHS at the dimer's start and end
    const cpx<double>* re0 = nullptr;
    const cpx<double>* ov = nullptr;   // FFT of this field's orientation block
    cpx<double>* k = nullptr;          // this field's spectrum, updated in place
    cpx<double>* torque = nullptr;     // k_sdd_torque: rs_i - re_i in, torque out
    const cpx<double>* re = nullptr;   //   re_i
    const double* sc = nullptr;
    double* part = nullptr;
};
// part[b] = sum w Re(work conj(ov)), work = alpha rs0 + (1 - alpha) re0
hipError_t launch_sdd_dot(const SddArgs& a, hipStream_t s);
// part[b] = sum w (2 dt D / N) |ov|^2 / (1 - dt D), part[kSddBlocks + b] = sum w (2 dt D / N) Re(conj(ov) d) / (1 - dt D)
// with d = k + dt (work - 2 ov sc[0])
hipError_t launch_sdd_sm(const SddArgs& a, hipStream_t s);
// k = (d - ov sc[2] / sc[1]) / (1 - dt D); part[b] = sum w |k_new - k_old|^2 / dt^2 / N,
// part[kSddBlocks + b] = max |k_new - k_old| / dt
hipError_t launch_sdd_update(const SddArgs& a, hipStream_t s);
// torque = torque - re - D ov len (torque holds rs on input); part[b] += sum w Re(torque conj(ov))
hipError_t launch_sdd_torque(const SddArgs& a, bool first, hipStream_t s);
// torque -= ov * sc[3]
hipError_t launch_sdd_project(const SddArgs& a, hipStream_t s);
// v -= c tau; part[b] = sum v_new^2, part[kSddBlocks + b] = max |tau|, part[2 kSddBlocks + b] = sum tau v_old
hipError_t launch_sdd_orient(double* v, const double* tau, long n, double c, double* part, hipStream_t s);
// single-thread bookkeeping of the SDD scalars and monitors (mode selects the step)
hipError_t launch_sdd_scalars(int mode, const double* part, double* sc, double* mon, double a0, double a1,
                              hipStream_t s);
template <typename T> hipError_t launch_etd_residual(const EtdArgs<T>& a, hipStream_t s);
template <typename T> hipError_t launch_kgrad(const KGradArgs<T>& a, int ndim, hipStream_t s);
template <typename T> hipError_t launch_kelastic(const KElasticArgs<T>& a, hipStream_t s);
template <typename T> hipError_t launch_kelastic_table(const KElasticArgs<T>& a, T* ktab, hipStream_t s);
// VolumeConservingLP.OnStepFinished (volumeConserving.go:31-50) on device state
// st = {Multiplier, CurrentIntegral, IsFirstUpdate}: the field integral is the sum of the
// partials px[0..nb), the indicator integral *ind (the zero mode of a k-space derived brick,
// or element 0 of a real field)
template <typename T> hipError_t launch_volcons_update(const double* px, long nb, const T* ind, double dt,
                                                       double* st, hipStream_t s);
// out = factor * sum_d w_d^2
template <typename T> hipError_t launch_sqsum(const T* const* w, int nw, T factor, T* out, long n, hipStream_t s);
template <typename T> hipError_t launch_copy(const T* src, T* dst, long n, hipStream_t s);
// per-block FP64 sums of f(x) = sum_q poly[q] x^q and of x over a real array
// (x is shifted by the FP64 offset before both sums)
template <typename T> hipError_t launch_real_sums(const T* x, long n, double offset, int npoly, const double* poly,
                                                  double* partial_f, double* partial_x, long nblocks, hipStream_t s);
// per-block FP64 sums of w |k|^2 |C|^2 over a half spectrum
template <typename T> hipError_t launch_kgrad_energy(const cpx<T>* C, KGrid g, long nmain, long nk,
                                                     double* partial, long nblocks, hipStream_t s);
// vectorised copy of n floats (n a multiple of 4): the bandwidth probe kernel
hipError_t launch_copy_bw(const float* src, float* dst, long n, hipStream_t s);
hipError_t launch_bw(int kind, const float* src, float* dst, long n, hipStream_t s);
// single-block reduction of partials; out[0] = sum a * ca + sum b * cb
hipError_t launch_reduce2(const double* a, long na, double ca, const double* b, long nb, double cb,
                          double* out, hipStream_t s);
// one entry of the energy series at the slot a device counter names (so that a captured
// graph can replay it): series[2 i] = sum bulk + ceg sum eg, series[2 i + 1] = sum mass,
// i = *counter, then *counter = i + 1
hipError_t launch_series_append(const double* bulk, long nbulk, const double* eg, long neg, double ceg,
                                const double* mass, double* series, long* counter, hipStream_t s);

}  // namespace gpf
