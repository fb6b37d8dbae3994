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
    double alpha = 1.0, beta = 0.0, gamma = 0.0;
    KPoly denom;
    const double* filter = nullptr;
    KGrid g;
    long nmain = 0, nk = 0;            // main elements; total (main + nyq)
};

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
};

template <typename T> hipError_t launch_derived(const DerivedArgs<T>& a, hipStream_t s);
template <typename T> hipError_t launch_kupdate(const KUpdateArgs<T>& a, hipStream_t s);
template <typename T> hipError_t launch_kgrad(const KGradArgs<T>& a, int ndim, hipStream_t s);
template <typename T> hipError_t launch_kelastic(const KElasticArgs<T>& a, hipStream_t s);
// out = factor * sum_d w_d^2
template <typename T> hipError_t launch_sqsum(const T* const* w, int nw, T factor, T* out, long n, hipStream_t s);
// indicator H(phi) = 3 phi^2 - 2 phi^3  (homoLinElast.go:10-12)
template <typename T> hipError_t launch_indicator(const T* phi, T* out, long n, hipStream_t s);
// P = H'(phi) (S - 2 e* H(phi))
template <typename T> hipError_t launch_elastic_real(const T* phi, const T* S, double two_estar, T* out, long n, hipStream_t s);
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
// single-block reduction of partials; out[0] = sum a * ca + sum b * cb
hipError_t launch_reduce2(const double* a, long na, double ca, const double* b, long nb, double cb,
                          double* out, hipStream_t s);

}  // namespace gpf
