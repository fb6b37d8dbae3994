/*
 * pf_oracle.c -- FP64 CPU restatement of gopf's spectral time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product
 * path is checked against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links,
 * loads or calls it.
 *
 * It restates, in plain C with complex128 arrays exactly as the reference
 * keeps them, the following reference code (paths relative to the gopf root):
 *   pfutil/fftWrap.go:16-95         FFTW C2C wrapper: FFT/IFFT/Freq/ConjugateNode
 *   pf/model.go:237-314             SyncDerivedFields, GetRHS, GetDenum
 *   pf/util.go:38-65                DerivedFieldCalcFromDesc (cmplx.Pow products)
 *   pf/rhsBuilder.go:125-242        ConcreteTerm / constructFunc term values
 *   pf/diffOp.go:25-29              LaplacianN.Eval
 *   pf/euler.go:16-47               Euler.Step
 *   pf/rk4.go:29-135                RK4.Step / Propagate and helpers
 *   pf/solver.go:70-85              Solver.Propagate (Step + OnStepFinished)
 *   pf/squareGradientTerm.go:38-65  SquaredGradient.Construct
 *   pf/homoLinElast.go:10-150       HomogeneousModulusLinElast
 *   elasticity/rank4.go:62-128      ContractLast, Isotropic, CubicMaterial
 *   elasticity/effectiveForce.go    EffectiveForce.Get
 *   elasticity/linearElasticity.go:16-98  DisplacementMatrixElement,
 *                                   Displacements, Strain, EnergyDensity
 *   pf/vandeven.go:13-40            NewVandeven / Eval
 *   pf/util.go:125-132              ApplyModalFilter
 *
 * The FFT that the reference obtains from FFTW3 (github.com/barnex/fftw,
 * go.mod:7) and gosfft v1.0.2 (go.mod:10) -- neither is vendored -- is
 * restated from the published DFT definition FFTW documents:
 *   forward  X[k] = sum_n x[n] exp(-2 pi i n k / N)   (sign -1, unnormalised)
 *   inverse  x[n] = sum_k X[k] exp(+2 pi i n k / N)   (sign +1, unnormalised)
 * as a mixed-radix Stockham transform (radices 4, 2, 3, 5 and a direct DFT
 * for other primes).  Multi-dimensional transforms use FFTW's row-major
 * layout: plan n = [D0, D1(, D2)] with the last index fastest.
 *
 * Go's complex division (Smith's algorithm, runtime/complex.go) and
 * cmplx.Pow (math/cmplx/pow.go) are restated so that the complex128
 * arithmetic follows the same formulas.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;

#define OR_PI 3.14159265358979323846

/* ------------------------------------------------------------------ */
/* Go complex arithmetic restatements                                  */
/* ------------------------------------------------------------------ */

/* Go runtime complex128div: Smith (1962) robust complex division. */
static zc go_cdiv(zc n, zc m)
{
    double a = creal(n), b = cimag(n), c = creal(m), d = cimag(m);
    double e, f;
    if (fabs(c) >= fabs(d)) {
        double ratio = d / c;
        double denom = c + ratio * d;
        e = (a + b * ratio) / denom;
        f = (b - a * ratio) / denom;
    } else {
        double ratio = c / d;
        double denom = d + ratio * c;
        e = (a * ratio + b) / denom;
        f = (b * ratio - a) / denom;
    }
    return e + I * f;
}

/* Go complex multiplication without fused operations. */
static inline zc go_cmul(zc x, zc y)
{
    double a = creal(x), b = cimag(x), c = creal(y), d = cimag(y);
    volatile double ac = a * c, bd = b * d, ad = a * d, bc = b * c;
    return (ac - bd) + I * (ad + bc);
}

/* math/cmplx Pow(x, complex(p, 0)) restated (real exponent only). */
static zc go_cpow(zc x, double p)
{
    if (creal(x) == 0.0 && cimag(x) == 0.0) {
        if (p == 0.0) return 1.0;
        if (p < 0.0) return INFINITY;
        return 0.0;
    }
    double mod = hypot(creal(x), cimag(x));
    if (mod == 0.0) return 0.0;
    double r = pow(mod, p);
    double theta = p * atan2(cimag(x), creal(x));
    return r * cos(theta) + I * (r * sin(theta));
}

/* ------------------------------------------------------------------ */
/* 1D mixed-radix Stockham FFT (FP64)                                  */
/* ------------------------------------------------------------------ */

typedef struct {
    int n;
    int nstages;
    int radix[32];
    int ns[32];         /* span before the stage */
    zc *tw[32];         /* per stage: ns * (radix - 1) twiddles, forward sign */
    zc *scratch;
} fft1d;

static int factorize(int n, int *out)
{
    int k = 0;
    while (n % 4 == 0) { out[k++] = 4; n /= 4; }
    while (n % 2 == 0) { out[k++] = 2; n /= 2; }
    while (n % 3 == 0) { out[k++] = 3; n /= 3; }
    while (n % 5 == 0) { out[k++] = 5; n /= 5; }
    for (int p = 7; n > 1; p += 2) {
        while (n % p == 0) { out[k++] = p; n /= p; }
    }
    return k;
}

static fft1d *fft1d_create(int n)
{
    fft1d *f = (fft1d *)calloc(1, sizeof(fft1d));
    f->n = n;
    f->nstages = (n > 1) ? factorize(n, f->radix) : 0;
    int ns = 1;
    for (int s = 0; s < f->nstages; ++s) {
        int r = f->radix[s];
        f->ns[s] = ns;
        f->tw[s] = (zc *)malloc(sizeof(zc) * (size_t)ns * (size_t)(r > 1 ? r - 1 : 1));
        for (int k = 0; k < ns; ++k)
            for (int m = 1; m < r; ++m) {
                /* exp(-2 pi i k m / (ns r)) computed from a reduced integer angle */
                long q = (long)k * m;
                long den = (long)ns * r;
                q %= den;
                double ang = -2.0 * OR_PI * (double)q / (double)den;
                f->tw[s][k * (r - 1) + (m - 1)] = cos(ang) + I * sin(ang);
            }
        ns *= r;
    }
    f->scratch = (zc *)malloc(sizeof(zc) * (size_t)(n > 0 ? n : 1));
    return f;
}

static void fft1d_destroy(fft1d *f)
{
    if (!f) return;
    for (int s = 0; s < f->nstages; ++s) free(f->tw[s]);
    free(f->scratch);
    free(f);
}

/* small DFT of radix r, sign sg (-1 forward, +1 inverse) */
static void dft_small(zc *a, int r, int sg)
{
    if (r == 2) {
        zc t0 = a[0] + a[1], t1 = a[0] - a[1];
        a[0] = t0; a[1] = t1;
    } else if (r == 4) {
        zc s02 = a[0] + a[2], d02 = a[0] - a[2];
        zc s13 = a[1] + a[3], d13 = a[1] - a[3];
        /* W4 = exp(sg * i pi/2) = sg * i */
        zc jd13 = (sg < 0) ? (cimag(d13) - I * creal(d13)) : (-cimag(d13) + I * creal(d13));
        a[0] = s02 + s13;
        a[1] = d02 + jd13;
        a[2] = s02 - s13;
        a[3] = d02 - jd13;
    } else {
        zc tmp[64];
        zc *b = tmp;
        zc *heap = NULL;
        if (r > 64) { heap = (zc *)malloc(sizeof(zc) * r); b = heap; }
        for (int q = 0; q < r; ++q) {
            zc acc = 0.0;
            for (int m = 0; m < r; ++m) {
                long e = ((long)m * q) % r;
                double ang = sg * 2.0 * OR_PI * (double)e / (double)r;
                acc += a[m] * (cos(ang) + I * sin(ang));
            }
            b[q] = acc;
        }
        for (int q = 0; q < r; ++q) a[q] = b[q];
        free(heap);
    }
}

/* in-place 1D transform of x (length n), sign sg */
static void fft1d_exec(fft1d *f, zc *x, int sg)
{
    int n = f->n;
    if (n <= 1) return;
    zc *src = x, *dst = f->scratch;
    zc a[64];
    for (int s = 0; s < f->nstages; ++s) {
        int r = f->radix[s], ns = f->ns[s];
        int nr = n / r;
        zc *tw = f->tw[s];
        zc *av = a;
        zc *heap = NULL;
        if (r > 64) { heap = (zc *)malloc(sizeof(zc) * r); av = heap; }
        for (int j = 0; j < nr; ++j) {
            int k = j % ns;
            for (int m = 0; m < r; ++m) av[m] = src[j + m * nr];
            if (ns > 1) {
                for (int m = 1; m < r; ++m) {
                    zc w = tw[k * (r - 1) + (m - 1)];
                    if (sg > 0) w = conj(w);
                    av[m] = av[m] * w;
                }
            }
            dft_small(av, r, sg);
            int idx = (j / ns) * ns * r + k;
            for (int m = 0; m < r; ++m) dst[idx + m * ns] = av[m];
        }
        free(heap);
        zc *t = src; src = dst; dst = t;
    }
    if (src != x) memcpy(x, src, sizeof(zc) * (size_t)n);
}

/* N-D plan over FFTW row-major layout n = [d0, d1(, d2)] */
typedef struct {
    int ndim;
    int dims[3];
    long total;
    fft1d *ax[3];
    zc *line;
} fftnd;

static fftnd *fftnd_create(int ndim, const int *dims)
{
    fftnd *p = (fftnd *)calloc(1, sizeof(fftnd));
    p->ndim = ndim;
    p->total = 1;
    int maxd = 1;
    for (int a = 0; a < ndim; ++a) {
        p->dims[a] = dims[a];
        p->total *= dims[a];
        p->ax[a] = fft1d_create(dims[a]);
        if (dims[a] > maxd) maxd = dims[a];
    }
    p->line = (zc *)malloc(sizeof(zc) * (size_t)maxd);
    return p;
}

static void fftnd_destroy(fftnd *p)
{
    if (!p) return;
    for (int a = 0; a < p->ndim; ++a) fft1d_destroy(p->ax[a]);
    free(p->line);
    free(p);
}

static void fftnd_exec(fftnd *p, zc *data, int sg)
{
    for (int a = p->ndim - 1; a >= 0; --a) {
        int n = p->dims[a];
        long stride = 1;
        for (int b = a + 1; b < p->ndim; ++b) stride *= p->dims[b];
        long outer = p->total / ((long)n * stride);
        for (long o = 0; o < outer; ++o) {
            for (long in = 0; in < stride; ++in) {
                zc *base = data + o * (long)n * stride + in;
                if (stride == 1) {
                    fft1d_exec(p->ax[a], base, sg);
                } else {
                    for (int e = 0; e < n; ++e) p->line[e] = base[(long)e * stride];
                    fft1d_exec(p->ax[a], p->line, sg);
                    for (int e = 0; e < n; ++e) base[(long)e * stride] = p->line[e];
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Freq / ConjugateNode  (pfutil/fftWrap.go:41-95)                     */
/* ------------------------------------------------------------------ */

static void ref_freq(int ndim, const int *dims, long i, double *res)
{
    long c = i % dims[1];
    long r = (i / dims[1]) % dims[0];
    res[1] = (double)c / (double)dims[1];
    res[0] = (double)r / (double)dims[0];
    if (ndim > 2) {
        long d = i / ((long)dims[0] * dims[1]);
        res[2] = (double)d / (double)dims[2];
    }
    for (int a = 0; a < ndim; ++a)
        if (res[a] > 0.5) res[a] -= 1.0;
}

long or_conjugate_node(int ndim, const int *dims, long i)
{
    long c = i % dims[1];
    long r = (i / dims[1]) % dims[0];
    long cc = (dims[1] - c) % dims[1];
    long cr = (dims[0] - r) % dims[0];
    long cd = 0;
    if (ndim == 3) {
        long d = i / ((long)dims[0] * dims[1]);
        cd = (dims[2] - d) % dims[2];
    }
    return cd * dims[0] * dims[1] + cr * dims[1] + cc;
}

void or_freq(int ndim, const int *dims, long i, double *out) { ref_freq(ndim, dims, i, out); }

/* floats.Norm(f, 2) */
static double norm2(const double *f, int n)
{
    double s = 0.0;
    for (int a = 0; a < n; ++a) s += f[a] * f[a];
    return sqrt(s);
}

/* LaplacianN.Eval factor: Pow(-Pow(2 pi |f|, 2), p)  (pf/diffOp.go:25-29) */
static double lap_factor(const double *f, int ndim, int p)
{
    double q = 2.0 * OR_PI * norm2(f, ndim);
    return pow(-pow(q, 2.0), (double)p);
}

/* ------------------------------------------------------------------ */
/* Stand-alone FFT entry (FourierTransform restatement)                */
/* ------------------------------------------------------------------ */

/* data: interleaved complex128, length prod(dims); dir -1 forward, +1 inverse (unnormalised) */
void or_fft(int ndim, const int *dims, double *data, int dir)
{
    fftnd *p = fftnd_create(ndim, dims);
    fftnd_exec(p, (zc *)data, dir < 0 ? -1 : 1);
    fftnd_destroy(p);
}

/* ------------------------------------------------------------------ */
/* Vandeven filter (pf/vandeven.go:13-40)                              */
/* ------------------------------------------------------------------ */

void or_vandeven_table(int order, double *data /* 1000 */)
{
    data[0] = 1.0;
    double prefactor = tgamma((double)(2 * order)) / (tgamma((double)order) * tgamma((double)order));
    double dx = 1.0 / 999.0;
    for (int i = 1; i < 1000; ++i) {
        double x = (double)i * dx;
        double i2 = pow(x * (1 - x), (double)(order - 1));
        double x1 = x - dx;
        double i1 = pow(x1 * (1.0 - x1), (double)(order - 1));
        data[i] = data[i - 1] - prefactor * 0.5 * (i1 + i2) * dx;
    }
}

double or_vandeven_eval(const double *data, double x)
{
    double N = 999.0;
    int idx = (int)(x * N);
    if (idx >= 999) return data[999];
    double dx = 1.0 / N;
    double dy = data[idx + 1] - data[idx];
    double x0 = (double)idx * dx;
    return data[idx] + (x - x0) * dy / dx;
}

/* ------------------------------------------------------------------ */
/* Elasticity helpers (elasticity/*.go)                                */
/* ------------------------------------------------------------------ */

static inline int r4(int i, int j, int k, int l) { return i * 27 + j * 9 + k * 3 + l; }

void or_cubic_material(double c11, double c12, double c44, double *t)
{
    memset(t, 0, sizeof(double) * 81);
    for (int i = 0; i < 3; ++i) t[r4(i, i, i, i)] = c11;
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j) {
            t[r4(i, i, j, j)] = c12;
            t[r4(j, j, i, i)] = c12;
            t[r4(i, j, i, j)] = c44;
            t[r4(j, i, j, i)] = c44;
        }
}

void or_isotropic(double bulk, double poisson, double *t)
{
    double shear = 3.0 * bulk * (1.0 - 2.0 * poisson) / (2.0 * (1.0 + poisson));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                    double v = 0.0;
                    if (i == j && k == l) v += bulk - 2.0 * shear / 3.0;
                    if (i == k && j == l) v += shear;
                    if (i == l && j == k) v += shear;
                    t[r4(i, j, k, l)] = v;
                }
}

static void contract_last(const double *C, const double *eps, double *out)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double cur = 0.0;
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) cur = cur + C[r4(i, j, k, l)] * eps[k * 3 + l];
            out[i * 3 + j] = cur;
        }
}

double or_energy_density(const double *C, const double *eps)
{
    double res = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) res += C[r4(i, j, k, l)] * eps[i * 3 + j] * eps[k * 3 + l];
    return 0.5 * res;
}

/* gonum Dense.Solve for a square system: LU with partial pivoting, 2 RHS columns */
static int solve3x2(double A[3][3], double B[3][2], double X[3][2])
{
    int piv[3] = {0, 1, 2};
    double a[3][3];
    memcpy(a, A, sizeof(a));
    for (int k = 0; k < 3; ++k) {
        int p = k;
        double best = fabs(a[k][k]);
        for (int i = k + 1; i < 3; ++i)
            if (fabs(a[i][k]) > best) { best = fabs(a[i][k]); p = i; }
        if (best == 0.0) return -1;
        if (p != k) {
            for (int j = 0; j < 3; ++j) { double t = a[k][j]; a[k][j] = a[p][j]; a[p][j] = t; }
            int t = piv[k]; piv[k] = piv[p]; piv[p] = t;
        }
        for (int i = k + 1; i < 3; ++i) {
            a[i][k] /= a[k][k];
            for (int j = k + 1; j < 3; ++j) a[i][j] -= a[i][k] * a[k][j];
        }
    }
    for (int c = 0; c < 2; ++c) {
        double y[3];
        for (int i = 0; i < 3; ++i) {
            double s = B[piv[i]][c];
            for (int j = 0; j < i; ++j) s -= a[i][j] * y[j];
            y[i] = s;
        }
        for (int i = 2; i >= 0; --i) {
            double s = y[i];
            for (int j = i + 1; j < 3; ++j) s -= a[i][j] * X[j][c];
            X[i][c] = s / a[i][i];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Model engine                                                        */
/* ------------------------------------------------------------------ */

enum { TERM_CONCRETE = 0, TERM_SQGRAD = 1, TERM_ELAST = 2, TERM_VOLCONS = 3, TERM_SVISC = 4, TERM_HESS = 5,
       TERM_CNOISE = 6 };

typedef struct {
    int kind;
    double coef_re, coef_im; /* concrete: sign * prod Pow(scalar, p); user: +-1 */
    int brick;               /* concrete: -1 none, else brick index */
    int has_lap;             /* concrete: term string contains LAP */
    int lap;                 /* concrete: LaplacianN power; user: total LAP prefix power */
    int user;                /* user term index (sqgrad or elastic table) */
} or_term;

typedef struct {
    int field_brick; /* brick holding the field u */
    double factor;
} or_sqgrad;

typedef struct {
    int field;        /* field index whose real part is lagged */
    double C[81];
    double misfit[9];
    double *lag;      /* h.Field, real, zero initialised (homoLinElast.go:145) */
    zc *inject_disp;  /* optional injected displacement (3 x n, component-major), test only */
} or_elast;

/* VolumeConservingLP (volumeConserving.go:8-17) */
typedef struct {
    int field_brick, ind_brick;
    double dt;
    long num_nodes;
    double multiplier, current_integral;
    int is_first_update;
} or_volcons;

/* SpectralViscosity (spectralViscosity.go:22-26) */
typedef struct {
    double eps, threshold;
    int power;
} or_svisc;

/* ConservativeNoise (noise.go:24-29): the bricks of its Dim current fields */
typedef struct {
    int dim;
    int bricks[3];
} or_cnoise;

/* TensorialHessian (tensorialHessian.go:33-36): K row-major, len 4 or 9 */
typedef struct {
    double K[9];
    int len;
} or_hess;

typedef struct {
    int ndim;
    int dims[3];
    long n;
    int nfields, nderived;
    zc **bricks; /* nfields + nderived arrays */
    /* derived field descriptors */
    int *dkind;        /* 0: product of complex pows (util.go:38-65); 1: real polynomial */
    int *dterm_start;  /* nderived + 1 */
    double *dterm_coef;/* per term */
    int *dfac_start;   /* nterms + 1 */
    int *dfac_field;
    double *dfac_pow;
    /* equations */
    int neq;
    int *eq_t_start, *eq_d_start; /* neq + 1 each */
    or_term *terms, *denum;
    /* user terms */
    int nsq, nel;
    or_sqgrad *sq;
    or_elast *el;
    int nvc, nsv, nhs, ncn;
    or_volcons *vc;
    or_svisc *sv;
    or_hess *hs;
    or_cnoise *cn;
    /* Gaussian-noise derived fields (kind 2): standard deviation, generator key, draws so far */
    double *dnoise_std;
    unsigned long long *dnoise_seed;
    unsigned *dcount;
    /* stepper */
    int stepper; /* 0 euler, 1 rk4 */
    double dt;
    long current_step;
    int filter_order;
    double vand[1000];
    fftnd *ft;
    zc *tmp, *tmp2, *work;
} or_model;

static void *dupmem(const void *p, size_t sz)
{
    void *q = malloc(sz ? sz : 1);
    if (sz) memcpy(q, p, sz);
    return q;
}

/*
 * Create the oracle model.  Arrays follow the layout documented in
 * oracle/engine.py (ModelSpec.flatten).
 */
void *or_model_create(int ndim, const int *dims, int nfields, int nderived,
                      const int *dkind, const int *dterm_start, const double *dterm_coef,
                      const int *dfac_start, const int *dfac_field, const double *dfac_pow,
                      int neq, const int *eq_t_start, const int *t_kind, const double *t_coef,
                      const int *t_brick, const int *t_has_lap, const int *t_lap, const int *t_user,
                      const int *eq_d_start, const int *d_kind, const double *d_coef,
                      const int *d_brick, const int *d_has_lap, const int *d_lap, const int *d_user,
                      int nsq, const int *sq_brick, const double *sq_factor,
                      int nel, const int *el_field, const double *el_C, const double *el_misfit,
                      int stepper, double dt, int filter_order)
{
    or_model *m = (or_model *)calloc(1, sizeof(or_model));
    m->ndim = ndim;
    m->n = 1;
    for (int a = 0; a < ndim; ++a) { m->dims[a] = dims[a]; m->n *= dims[a]; }
    m->nfields = nfields;
    m->nderived = nderived;
    int nb = nfields + nderived;
    m->bricks = (zc **)calloc((size_t)nb, sizeof(zc *));
    for (int b = 0; b < nb; ++b) m->bricks[b] = (zc *)calloc((size_t)m->n, sizeof(zc));
    m->dkind = (int *)dupmem(dkind, sizeof(int) * nderived);
    m->dnoise_std = (double *)calloc((size_t)(nderived ? nderived : 1), sizeof(double));
    m->dnoise_seed = (unsigned long long *)calloc((size_t)(nderived ? nderived : 1), sizeof(unsigned long long));
    m->dcount = (unsigned *)calloc((size_t)(nderived ? nderived : 1), sizeof(unsigned));
    m->dterm_start = (int *)dupmem(dterm_start, sizeof(int) * (nderived + 1));
    int nt = dterm_start[nderived];
    m->dterm_coef = (double *)dupmem(dterm_coef, sizeof(double) * nt);
    m->dfac_start = (int *)dupmem(dfac_start, sizeof(int) * (nt + 1));
    int nfac = dfac_start[nt];
    m->dfac_field = (int *)dupmem(dfac_field, sizeof(int) * nfac);
    m->dfac_pow = (double *)dupmem(dfac_pow, sizeof(double) * nfac);
    m->neq = neq;
    m->eq_t_start = (int *)dupmem(eq_t_start, sizeof(int) * (neq + 1));
    m->eq_d_start = (int *)dupmem(eq_d_start, sizeof(int) * (neq + 1));
    int ntt = eq_t_start[neq], ndd = eq_d_start[neq];
    m->terms = (or_term *)calloc((size_t)(ntt ? ntt : 1), sizeof(or_term));
    m->denum = (or_term *)calloc((size_t)(ndd ? ndd : 1), sizeof(or_term));
    for (int t = 0; t < ntt; ++t) {
        or_term *x = &m->terms[t];
        x->kind = t_kind[t]; x->coef_re = t_coef[2 * t]; x->coef_im = t_coef[2 * t + 1];
        x->brick = t_brick[t]; x->has_lap = t_has_lap[t]; x->lap = t_lap[t]; x->user = t_user[t];
    }
    for (int t = 0; t < ndd; ++t) {
        or_term *x = &m->denum[t];
        x->kind = d_kind[t]; x->coef_re = d_coef[2 * t]; x->coef_im = d_coef[2 * t + 1];
        x->brick = d_brick[t]; x->has_lap = d_has_lap[t]; x->lap = d_lap[t]; x->user = d_user[t];
    }
    m->nsq = nsq;
    m->sq = (or_sqgrad *)calloc((size_t)(nsq ? nsq : 1), sizeof(or_sqgrad));
    for (int s = 0; s < nsq; ++s) { m->sq[s].field_brick = sq_brick[s]; m->sq[s].factor = sq_factor[s]; }
    m->nel = nel;
    m->el = (or_elast *)calloc((size_t)(nel ? nel : 1), sizeof(or_elast));
    for (int e = 0; e < nel; ++e) {
        m->el[e].field = el_field[e];
        memcpy(m->el[e].C, el_C + 81 * e, sizeof(double) * 81);
        memcpy(m->el[e].misfit, el_misfit + 9 * e, sizeof(double) * 9);
        m->el[e].lag = (double *)calloc((size_t)m->n, sizeof(double));
    }
    m->stepper = stepper;
    m->dt = dt;
    m->filter_order = filter_order;
    if (filter_order > 0) or_vandeven_table(filter_order, m->vand);
    m->ft = fftnd_create(ndim, dims);
    m->tmp = (zc *)calloc((size_t)m->n, sizeof(zc));
    m->tmp2 = (zc *)calloc((size_t)m->n, sizeof(zc));
    m->work = (zc *)calloc((size_t)m->n, sizeof(zc));
    return m;
}

void or_model_destroy(void *h)
{
    or_model *m = (or_model *)h;
    if (!m) return;
    for (int b = 0; b < m->nfields + m->nderived; ++b) free(m->bricks[b]);
    free(m->bricks);
    free(m->dkind); free(m->dterm_start); free(m->dterm_coef); free(m->dfac_start);
    free(m->dfac_field); free(m->dfac_pow);
    free(m->eq_t_start); free(m->eq_d_start); free(m->terms); free(m->denum);
    free(m->sq);
    for (int e = 0; e < m->nel; ++e) { free(m->el[e].lag); free(m->el[e].inject_disp); }
    free(m->el);
    free(m->vc); free(m->sv); free(m->hs); free(m->cn);
    free(m->dnoise_std); free(m->dnoise_seed); free(m->dcount);
    fftnd_destroy(m->ft);
    free(m->tmp); free(m->tmp2); free(m->work);
    free(m);
}

void or_set_brick(void *h, int b, const double *data)
{
    or_model *m = (or_model *)h;
    memcpy(m->bricks[b], data, sizeof(zc) * (size_t)m->n);
}

void or_get_brick(void *h, int b, double *out)
{
    or_model *m = (or_model *)h;
    memcpy(out, m->bricks[b], sizeof(zc) * (size_t)m->n);
}

void or_set_lag(void *h, int e, const double *data)
{
    or_model *m = (or_model *)h;
    memcpy(m->el[e].lag, data, sizeof(double) * (size_t)m->n);
}

void or_inject_displacement(void *h, int e, const double *disp /* 3 x n interleaved complex */)
{
    or_model *m = (or_model *)h;
    if (!m->el[e].inject_disp) m->el[e].inject_disp = (zc *)malloc(sizeof(zc) * 3 * (size_t)m->n);
    memcpy(m->el[e].inject_disp, disp, sizeof(zc) * 3 * (size_t)m->n);
}

long or_current_step(void *h) { return ((or_model *)h)->current_step; }
void or_set_current_step(void *h, long s) { ((or_model *)h)->current_step = s; }

/* Model.SyncDerivedFields (model.go:237-241) */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11): ten rounds of two 32x32->64 multiplies with the Weyl key schedule */
void or_philox(const unsigned *ctr, const unsigned *key, unsigned *out)
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* the normal draw of node i: Philox on (i, counter, stream) keyed by seed, two 53-bit
 * uniforms in (0, 1], Box-Muller */
double or_noise_normal(long i, unsigned counter, unsigned stream, unsigned long long seed)
{
    unsigned c[4] = {(unsigned)((unsigned long long)i & 0xffffffffu), (unsigned)((unsigned long long)i >> 32),
                     counter, stream};
    unsigned k[2] = {(unsigned)(seed & 0xffffffffu), (unsigned)(seed >> 32)}, o[4];
    or_philox(c, k, o);
    double s53 = 1.0 / 9007199254740992.0;
    double u1 = ((double)((((unsigned long long)(o[0] >> 5)) << 26) | (o[1] >> 6)) + 0.5) * s53;
    double u2 = ((double)((((unsigned long long)(o[2] >> 5)) << 26) | (o[3] >> 6)) + 0.5) * s53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

void or_set_noise(void *h, int d, double std, unsigned long long seed)
{
    or_model *m = (or_model *)h;
    m->dnoise_std[d] = std;
    m->dnoise_seed[d] = seed;
}

/* ConservativeNoise + RegisterExplicitTerm, in the order of the model's user table */
void or_add_cnoise(void *h, int dim, const int *bricks)
{
    or_model *m = (or_model *)h;
    m->cn = (or_cnoise *)realloc(m->cn, sizeof(or_cnoise) * (size_t)(m->ncn + 1));
    or_cnoise *v = &m->cn[m->ncn++];
    v->dim = dim;
    for (int c = 0; c < 3; ++c) v->bricks[c] = (c < dim) ? bricks[c] : -1;
}

static void sync_derived(or_model *m)
{
    for (int d = 0; d < m->nderived; ++d) {
        zc *out = m->bricks[m->nfields + d];
        int t0 = m->dterm_start[d], t1 = m->dterm_start[d + 1];
        unsigned draw = m->dcount[d]++;
        if (m->dkind[d] == 2) {
            /* WhiteNoise.Generate / the current fields of ConservativeNoise (noise.go:19-22, 90-97) */
            for (long i = 0; i < m->n; ++i)
                out[i] = m->dnoise_std[d] * or_noise_normal(i, draw, (unsigned)d, m->dnoise_seed[d]);
            continue;
        }
        if (m->dkind[d] == 0) {
            /* DerivedFieldCalcFromDesc: data = 1; data *= cmplx.Pow(field, p) in order */
            int f0 = m->dfac_start[t0], f1 = m->dfac_start[t0 + 1];
            for (long i = 0; i < m->n; ++i) {
                zc v = 1.0;
                for (int f = f0; f < f1; ++f)
                    v = go_cmul(v, go_cpow(m->bricks[m->dfac_field[f]][i], m->dfac_pow[f]));
                out[i] = v;
            }
        } else {
            /* real polynomial of the real parts of the fields */
            for (long i = 0; i < m->n; ++i) {
                double acc = 0.0;
                for (int t = t0; t < t1; ++t) {
                    double v = m->dterm_coef[t];
                    for (int f = m->dfac_start[t]; f < m->dfac_start[t + 1]; ++f) {
                        double x = creal(m->bricks[m->dfac_field[f]][i]);
                        double p = m->dfac_pow[f];
                        double xp;
                        if (p == (double)(int)p && fabs(p) <= 16) {
                            int ip = (int)fabs(p);
                            xp = 1.0;
                            for (int q = 0; q < ip; ++q) xp *= x;
                            if (p < 0) xp = 1.0 / xp;
                        } else {
                            xp = pow(x, p);
                        }
                        v *= xp;
                    }
                    acc += v;
                }
                out[i] = acc;
            }
        }
    }
}

/* SquaredGradient.Construct (squareGradientTerm.go:38-65) */
static void eval_sqgrad(or_model *m, const or_sqgrad *s, zc *field)
{
    long n = m->n;
    int dim = m->ndim;
    zc *work = m->work;
    for (long i = 0; i < n; ++i) field[i] = 0.0;
    const zc *u = m->bricks[s->field_brick];
    double f[3];
    for (int d = 0; d < dim; ++d) {
        for (long i = 0; i < n; ++i) {
            ref_freq(m->ndim, m->dims, i, f);
            if (fabs(f[d] - 0.5) < 1e-10) f[d] = 0.0;
            work[i] = go_cmul(u[i], 0.0 + I * (2.0 * OR_PI * f[d]));
        }
        fftnd_exec(m->ft, work, +1);
        for (long i = 0; i < n; ++i) work[i] = go_cpow(go_cdiv(work[i], (double)n), 2.0);
        fftnd_exec(m->ft, work, -1);
        for (long i = 0; i < n; ++i) field[i] += go_cmul(work[i], s->factor);
    }
}

/* HomogeneousModulusLinElast.Construct (homoLinElast.go:47-99) */
static void eval_elast(or_model *m, const or_elast *e, zc *field)
{
    long n = m->n;
    int Dim = m->ndim;
    for (long i = 0; i < n; ++i) field[i] = 0.0;
    zc *work = (zc *)malloc(sizeof(zc) * n);
    for (long i = 0; i < n; ++i) {
        double x = e->lag[i];
        work[i] = 3.0 * x * x - 2.0 * x * x * x;
    }
    fftnd_exec(m->ft, work, -1);
    double sig[9];
    contract_last(e->C, e->misfit, sig); /* EffStress = C : misfit */
    zc *disp;
    int own_disp = 0;
    if (e->inject_disp) {
        disp = e->inject_disp;
    } else {
        own_disp = 1;
        disp = (zc *)calloc((size_t)(3 * n), sizeof(zc));
        for (long i = 0; i < n; ++i) {
            double k[3] = {0.0, 0.0, 0.0};
            ref_freq(m->ndim, m->dims, i, k); /* h.Freq pads to length 3 */
            /* Force: components c < Dim (effectiveForce.go:26-35) */
            zc F[3] = {0.0, 0.0, 0.0};
            for (int c = 0; c < Dim; ++c) {
                zc acc = 0.0;
                for (int j = 0; j < 3; ++j)
                    acc += go_cmul(0.0 + I * (-sig[c * 3 + j] * 2.0 * OR_PI * k[j]), work[i]);
                F[c] = acc;
            }
            /* Displacements (linearElasticity.go:33-63) */
            if (fabs(k[0]) < 1e-10 && fabs(k[1]) < 1e-10 && fabs(k[2]) < 1e-10) continue;
            double A[3][3], B[3][2], X[3][2];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    double el = 0.0;
                    for (int j = 0; j < 3; ++j)
                        for (int l = 0; l < 3; ++l) el += e->C[r4(a, j, b, l)] * k[j] * k[l];
                    A[a][b] = el * pow(2.0 * OR_PI, 2);
                }
            for (int a = 0; a < 3; ++a) { B[a][0] = creal(F[a]); B[a][1] = cimag(F[a]); }
            if (solve3x2(A, B, X) == 0)
                for (int a = 0; a < 3; ++a) disp[a * n + i] = X[a][0] + I * X[a][1];
        }
    }
    /* work <- IndicatorDeriv(h.Field) */
    for (long i = 0; i < n; ++i) {
        double x = e->lag[i];
        work[i] = 6.0 * x - 6.0 * x * x;
    }
    zc *strain = (zc *)malloc(sizeof(zc) * n);
    for (int a = 0; a < Dim; ++a)
        for (int b = a; b < Dim; ++b) {
            /* Strain (linearElasticity.go:67-83) */
            for (long i = 0; i < n; ++i) {
                double k[3] = {0.0, 0.0, 0.0};
                ref_freq(m->ndim, m->dims, i, k);
                double fm = k[a], fn = k[b];
                if (fabs(fabs(fm) - 0.5) < 1e-10) fm = 0.0;
                if (fabs(fabs(fn) - 0.5) < 1e-10) fn = 0.0;
                strain[i] = go_cmul(0.0 + I * (OR_PI * fn), disp[a * n + i]) +
                            go_cmul(0.0 + I * (OR_PI * fm), disp[b * n + i]);
            }
            fftnd_exec(m->ft, strain, +1);
            for (long i = 0; i < n; ++i) strain[i] = go_cmul(go_cdiv(strain[i], (double)n), work[i]);
            fftnd_exec(m->ft, strain, -1);
            double factor = (a != b) ? 2.0 : 1.0;
            for (long i = 0; i < n; ++i) field[i] += go_cmul(factor * sig[a * 3 + b], strain[i]);
        }
    double eden = or_energy_density(e->C, e->misfit);
    for (long i = 0; i < n; ++i) {
        double x = e->lag[i];
        work[i] = (3.0 * x * x - 2.0 * x * x * x) * (6.0 * x - 6.0 * x * x);
    }
    fftnd_exec(m->ft, work, -1);
    for (long i = 0; i < n; ++i) field[i] -= go_cmul(2.0 * eden, work[i]);
    free(strain);
    free(work);
    if (own_disp) free(disp);
}

/* NewVolumeConservingLP + RegisterExplicitTerm, in the order of the model's user table */
void or_add_volcons(void *h, int field_brick, int ind_brick, double dt, long num_nodes)
{
    or_model *m = (or_model *)h;
    m->vc = (or_volcons *)realloc(m->vc, sizeof(or_volcons) * (size_t)(m->nvc + 1));
    or_volcons *v = &m->vc[m->nvc++];
    v->field_brick = field_brick; v->ind_brick = ind_brick; v->dt = dt; v->num_nodes = num_nodes;
    v->multiplier = 0.0; v->current_integral = 0.0; v->is_first_update = 1;
}

/* SpectralViscosity + RegisterImplicitTerm, in the order of the model's user table */
void or_add_svisc(void *h, double eps, double threshold, int power)
{
    or_model *m = (or_model *)h;
    m->sv = (or_svisc *)realloc(m->sv, sizeof(or_svisc) * (size_t)(m->nsv + 1));
    or_svisc *v = &m->sv[m->nsv++];
    v->eps = eps; v->threshold = threshold; v->power = power;
}

double or_volcons_multiplier(void *h, int v) { return ((or_model *)h)->vc[v].multiplier; }

/* TensorialHessian + RegisterImplicitTerm, in the order of the model's user table */
void or_add_hessian(void *h, const double *K, int len)
{
    or_model *m = (or_model *)h;
    m->hs = (or_hess *)realloc(m->hs, sizeof(or_hess) * (size_t)(m->nhs + 1));
    or_hess *v = &m->hs[m->nhs++];
    memcpy(v->K, K, sizeof(double) * 9);
    v->len = len;
}

/* interpolant (spectralViscosity.go:29-39) */
double or_sv_interpolant(double f, double peak)
{
    double frac = 1.0 / 3.0;
    if (f < frac * peak) return 0.0;
    else if (f > peak) return 1.0;
    double x = 1.5 * (f - frac * peak) / peak;
    return 2.0 * x * x - 3.0 * x * x * x;
}

/* math.Pow(x, n), n a non-negative integer: repeated squaring */
static double go_pow_int(double x, int n)
{
    double a = 1.0, b = x;
    for (int i = n; i != 0; i >>= 1) {
        if (i & 1) a *= b;
        b *= b;
    }
    return a;
}

/* evaluate one term (closure) into out */
static void eval_term(or_model *m, const or_term *t, zc *out)
{
    long n = m->n;
    double f[3];
    if (t->kind == TERM_CONCRETE) {
        zc coef = t->coef_re + I * t->coef_im;
        const zc *br = (t->brick >= 0) ? m->bricks[t->brick] : NULL;
        for (long i = 0; i < n; ++i) {
            zc v = coef;
            if (br) v = go_cmul(v, br[i]);
            if (t->has_lap) {
                ref_freq(m->ndim, m->dims, i, f);
                v = go_cmul(v, lap_factor(f, m->ndim, t->lap));
            }
            out[i] = v;
        }
        return;
    }
    if (t->kind == TERM_SQGRAD) {
        eval_sqgrad(m, &m->sq[t->user], out);
    } else if (t->kind == TERM_VOLCONS) {
        /* VolumeConservingLP.Construct (volumeConserving.go:19-28) */
        const or_volcons *v = &m->vc[t->user];
        const zc *ind = m->bricks[v->ind_brick];
        for (long i = 0; i < n; ++i) out[i] = go_cmul(ind[i], v->multiplier + 0.0 * I);
    } else if (t->kind == TERM_SVISC) {
        /* SpectralViscosity.Construct (spectralViscosity.go:44-52) */
        const or_svisc *v = &m->sv[t->user];
        for (long i = 0; i < n; ++i) {
            ref_freq(m->ndim, m->dims, i, f);
            double d = 0.0;
            for (int a = 0; a < m->ndim; ++a) d += f[a] * f[a];
            double fr = sqrt(d);
            out[i] = -v->eps * or_sv_interpolant(fr, v->threshold) * go_pow_int(fr, v->power);
        }
    } else if (t->kind == TERM_CNOISE) {
        /* ConservativeNoise.Construct (noise.go:60-78) */
        const or_cnoise *v = &m->cn[t->user];
        for (long i = 0; i < n; ++i) {
            ref_freq(m->ndim, m->dims, i, f);
            zc acc = 0.0;
            for (int c = 0; c < v->dim; ++c) {
                double fc = f[c];
                if (fabs(fabs(fc) - 0.5) > 1e-6) {
                    double omega_half = OR_PI * fc;
                    acc += go_cmul(0.0 + 2.0 * sin(omega_half) * I, m->bricks[v->bricks[c]][i]);
                }
            }
            out[i] = acc;
        }
    } else if (t->kind == TERM_HESS) {
        /* TensorialHessian.Construct (tensorialHessian.go:37-60); GetCoeff's d is 3 only for len 9 */
        const or_hess *v = &m->hs[t->user];
        int d = (v->len == 9) ? 3 : 2;
        for (long i = 0; i < n; ++i) {
            ref_freq(m->ndim, m->dims, i, f);
            zc acc = 0.0;
            for (int j = 0; j < m->ndim; ++j) acc += -4.0 * OR_PI * OR_PI * f[j] * f[j] * v->K[j * d + j];
            for (int j = 0; j < m->ndim; ++j)
                for (int k = j + 1; k < m->ndim; ++k) acc += -8.0 * OR_PI * OR_PI * f[j] * f[k] * v->K[j * d + k];
            out[i] = acc;
        }
    } else {
        eval_elast(m, &m->el[t->user], out);
    }
    /* constructFunc prefixes: sign flips and LAP powers */
    for (long i = 0; i < n; ++i) {
        zc v = out[i];
        if (t->lap > 0) {
            ref_freq(m->ndim, m->dims, i, f);
            v = go_cmul(v, lap_factor(f, m->ndim, t->lap));
        }
        if (t->coef_re < 0) v = -v;
        out[i] = v;
    }
}

/* Model.GetRHS (model.go:283-303) */
static void get_rhs(or_model *m, int eq, zc *data)
{
    long n = m->n;
    for (long i = 0; i < n; ++i) data[i] = 0.0;
    for (int t = m->eq_t_start[eq]; t < m->eq_t_start[eq + 1]; ++t) {
        eval_term(m, &m->terms[t], m->tmp2);
        for (long i = 0; i < n; ++i) data[i] += m->tmp2[i];
    }
}

/* Model.GetDenum (model.go:306-314) */
static void get_denum(or_model *m, int eq, zc *data)
{
    long n = m->n;
    for (long i = 0; i < n; ++i) data[i] = 0.0;
    for (int t = m->eq_d_start[eq]; t < m->eq_d_start[eq + 1]; ++t) {
        eval_term(m, &m->denum[t], m->tmp2);
        for (long i = 0; i < n; ++i) data[i] += m->tmp2[i];
    }
}

void or_get_rhs(void *h, int eq, double *out) { get_rhs((or_model *)h, eq, (zc *)out); }
void or_get_denum(void *h, int eq, double *out) { get_denum((or_model *)h, eq, (zc *)out); }
void or_sync_derived(void *h) { sync_derived((or_model *)h); }
void or_fft_brick(void *h, int b, int dir) { or_model *m = (or_model *)h; fftnd_exec(m->ft, m->bricks[b], dir < 0 ? -1 : 1); }

/* ApplyModalFilter (util.go:125-132) */
static void apply_filter(or_model *m, zc *d)
{
    double f[3];
    for (long i = 0; i < m->n; ++i) {
        ref_freq(m->ndim, m->dims, i, f);
        double frad = sqrt(f[0] * f[0] + f[1] * f[1] + (m->ndim > 2 ? f[2] * f[2] : 0.0));
        double value = frad * 2.0 / OR_PI;
        d[i] = go_cmul(d[i], or_vandeven_eval(m->vand, value));
    }
}

static void div_real(zc *d, long n, double fac)
{
    for (long i = 0; i < n; ++i) d[i] = go_cdiv(d[i], fac);
}

/* Euler.Step (euler.go:16-47) */
static void euler_step(or_model *m)
{
    long n = m->n;
    double dt = m->dt;
    sync_derived(m);
    for (int f = 0; f < m->nfields; ++f) fftnd_exec(m->ft, m->bricks[f], -1);
    for (int d = 0; d < m->nderived; ++d) fftnd_exec(m->ft, m->bricks[m->nfields + d], -1);
    zc *rhs = (zc *)malloc(sizeof(zc) * n);
    zc *den = (zc *)malloc(sizeof(zc) * n);
    for (int i = 0; i < m->nfields; ++i) {
        get_rhs(m, i, rhs);
        get_denum(m, i, den);
        zc *d = m->bricks[i];
        for (long j = 0; j < n; ++j)
            d[j] = go_cdiv(d[j] + go_cmul(dt, rhs[j]), 1.0 - go_cmul(dt, den[j]));
        if (m->filter_order > 0) apply_filter(m, d);
    }
    free(rhs);
    free(den);
    for (int f = 0; f < m->nfields; ++f) {
        fftnd_exec(m->ft, m->bricks[f], +1);
        div_real(m->bricks[f], n, (double)n);
    }
    m->current_step++;
}

/* RK4.Step (rk4.go:29-127) */
static void rk4_step(or_model *m)
{
    long n = m->n;
    int nf = m->nfields;
    double dt = m->dt;
    sync_derived(m);
    for (int f = 0; f < nf; ++f) fftnd_exec(m->ft, m->bricks[f], -1);
    zc **initial = (zc **)malloc(sizeof(zc *) * nf);
    zc **final = (zc **)malloc(sizeof(zc *) * nf);
    zc **kf = (zc **)malloc(sizeof(zc *) * nf);
    for (int f = 0; f < nf; ++f) {
        initial[f] = (zc *)dupmem(m->bricks[f], sizeof(zc) * n);
        final[f] = (zc *)dupmem(m->bricks[f], sizeof(zc) * n);
        kf[f] = (zc *)dupmem(m->bricks[f], sizeof(zc) * n);
    }
    zc *den = (zc *)malloc(sizeof(zc) * n);
    /* firstCorrection */
    for (int d = 0; d < m->nderived; ++d) fftnd_exec(m->ft, m->bricks[nf + d], -1);
    for (int f = 0; f < nf; ++f) get_rhs(m, f, kf[f]);
    const double fac_prep[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const double fac_corr[3] = {0.5, 0.5, 1.0};
    for (int stage = 0; stage < 4; ++stage) {
        /* PrepareNextCorrection */
        for (int f = 0; f < nf; ++f) {
            zc c = fac_prep[stage] * dt;
            for (long j = 0; j < n; ++j) final[f][j] += go_cmul(c, kf[f][j]);
            memcpy(m->bricks[f], initial[f], sizeof(zc) * n);
        }
        if (stage == 3) break;
        /* correction(factor) */
        double factor = fac_corr[stage];
        for (int f = 0; f < nf; ++f) {
            get_denum(m, f, den);
            zc *d = m->bricks[f];
            zc c = factor * dt;
            for (long j = 0; j < n; ++j) {
                d[j] += go_cmul(c, kf[f][j]);
                d[j] = go_cdiv(d[j], 1.0 - go_cmul(c, den[j]));
            }
            fftnd_exec(m->ft, d, +1);
            div_real(d, n, (double)n);
        }
        sync_derived(m);
        for (int f = 0; f < nf; ++f) fftnd_exec(m->ft, m->bricks[f], -1);
        for (int d = 0; d < m->nderived; ++d) fftnd_exec(m->ft, m->bricks[nf + d], -1);
        for (int f = 0; f < nf; ++f) get_rhs(m, f, kf[f]);
    }
    for (int f = 0; f < nf; ++f) {
        get_denum(m, f, den);
        for (long j = 0; j < n; ++j) final[f][j] = go_cdiv(final[f][j], 1.0 - go_cmul(dt, den[j]));
        memcpy(m->bricks[f], final[f], sizeof(zc) * n);
        if (m->filter_order > 0) apply_filter(m, m->bricks[f]);
    }
    for (int f = 0; f < nf; ++f) {
        fftnd_exec(m->ft, m->bricks[f], +1);
        div_real(m->bricks[f], n, (double)n);
    }
    for (int f = 0; f < nf; ++f) { free(initial[f]); free(final[f]); free(kf[f]); }
    free(initial); free(final); free(kf); free(den);
}

/* TimeStepper.Step */
void or_step(void *h)
{
    or_model *m = (or_model *)h;
    if (m->stepper == 0) euler_step(m);
    else rk4_step(m);
}

/* Euler.Propagate / RK4.Propagate (euler.go:55-59, rk4.go:130-135) */
void or_stepper_propagate(void *h, int nsteps)
{
    or_model *m = (or_model *)h;
    for (int s = 0; s < nsteps; ++s) {
        or_step(m);
        if (m->stepper == 1) m->current_step++;
    }
}

double or_get_time(void *h)
{
    or_model *m = (or_model *)h;
    return (double)m->current_step * m->dt;
}

/* OnStepFinished of every term with state: HomogeneousModulusLinElast's lagged field
 * (homoLinElast.go:130-134) and VolumeConservingLP's multiplier (volumeConserving.go:31-50;
 * the field brick is real space after a step, a derived indicator brick still k-space) */
void or_on_step_finished(void *h)
{
    or_model *m = (or_model *)h;
    for (int e = 0; e < m->nel; ++e) {
        zc *f = m->bricks[m->el[e].field];
        for (long i = 0; i < m->n; ++i) m->el[e].lag[i] = creal(f[i]);
    }
    for (int v = 0; v < m->nvc; ++v) {
        or_volcons *c = &m->vc[v];
        double field_integral = 0.0;
        for (long i = 0; i < c->num_nodes; ++i) field_integral += creal(m->bricks[c->field_brick][i]);
        double indicator_integral = creal(m->bricks[c->ind_brick][0]);
        if (c->is_first_update) {
            c->current_integral = field_integral;
            c->is_first_update = 0;
        } else {
            double delta = field_integral - c->current_integral;
            c->current_integral = field_integral;
            c->multiplier = c->multiplier - delta / (c->dt * indicator_integral);
        }
    }
}

/* Solver.Propagate (solver.go:70-85): Step, then OnStepFinished of every term. */
void or_solver_propagate(void *h, int nsteps)
{
    or_model *m = (or_model *)h;
    for (int s = 0; s < nsteps; ++s) {
        or_step(m);
        or_on_step_finished(m);
    }
}

/* Evaluate a single user term's closure value (for the term known-answer tests). */
void or_eval_user_term(void *h, int kind, int user, double *out)
{
    or_model *m = (or_model *)h;
    or_term t = {0};
    t.kind = kind;
    t.coef_re = 1.0;
    t.user = user;
    if (kind == TERM_ELAST) eval_elast(m, &m->el[user], (zc *)out);
    else eval_term(m, &t, (zc *)out);
}

/* ------------------------------------------------------------------ */
/* Cahn-Hilliard free energy (no reference counterpart)                */
/*   F = sum_x f(c) + (kappa/2) sum_x |grad c|^2                        */
/*   f(c) = sum_p a_p c^p; gradient term from Parseval in k-space.      */
/* ------------------------------------------------------------------ */
double or_ch_energy(int ndim, const int *dims, const double *c, int npoly, const double *poly,
                    double kappa)
{
    long n = 1;
    for (int a = 0; a < ndim; ++a) n *= dims[a];
    double bulk = 0.0, comp = 0.0;
    zc *w = (zc *)malloc(sizeof(zc) * n);
    for (long i = 0; i < n; ++i) {
        double x = c[i], v = 0.0, xp = 1.0;
        for (int p = 0; p < npoly; ++p) { v += poly[p] * xp; xp *= x; }
        /* Neumaier compensated sum */
        double t = bulk + v;
        if (fabs(bulk) >= fabs(v)) comp += (bulk - t) + v;
        else comp += (v - t) + bulk;
        bulk = t;
        w[i] = x;
    }
    bulk += comp;
    fftnd *p = fftnd_create(ndim, dims);
    fftnd_exec(p, w, -1);
    double g = 0.0, gc = 0.0, f[3];
    for (long i = 0; i < n; ++i) {
        ref_freq(ndim, dims, i, f);
        double k2 = 0.0;
        for (int a = 0; a < ndim; ++a) k2 += f[a] * f[a];
        double v = 4.0 * OR_PI * OR_PI * k2 * (creal(w[i]) * creal(w[i]) + cimag(w[i]) * cimag(w[i]));
        double t = g + v;
        if (fabs(g) >= fabs(v)) gc += (g - t) + v;
        else gc += (v - t) + g;
        g = t;
    }
    g += gc;
    fftnd_destroy(p);
    free(w);
    return bulk + 0.5 * kappa * g / (double)n;
}

/* ------------------------------------------------------------------ */
/* Reverse mode of Euler.Step (no reference counterpart)               */
/* ------------------------------------------------------------------ */
/*
 * d ln b / d c of a derived brick, times b: the derivative of a derived field of
 * field 0 at the value x (product of powers, util.go:38-65, or a real polynomial).
 */
static double ipow_real(double x, double p)
{
    if (p == (double)(int)p && fabs(p) <= 16) {
        int ip = (int)fabs(p);
        double v = 1.0;
        for (int q = 0; q < ip; ++q) v *= x;
        return (p < 0) ? 1.0 / v : v;
    }
    return pow(x, p);
}

static double derived_deriv(const or_model *m, int d, double x)
{
    double acc = 0.0;
    int t0 = m->dterm_start[d], t1 = m->dterm_start[d + 1];
    if (m->dkind[d] == 0) t1 = t0 + 1; /* one product term */
    for (int t = t0; t < t1; ++t) {
        double coef = (m->dkind[d] == 0) ? 1.0 : m->dterm_coef[t];
        int f0 = m->dfac_start[t], f1 = m->dfac_start[t + 1];
        for (int j = f0; j < f1; ++j) {
            double v = coef * m->dfac_pow[j] * ipow_real(x, m->dfac_pow[j] - 1.0);
            for (int i = f0; i < f1; ++i)
                if (i != j) v *= ipow_real(x, m->dfac_pow[i]);
            acc += v;
        }
    }
    return acc;
}

/*
 * Gradient of J = sum_x (c_T(x) - target(x))^2 after nsteps Euler steps
 * (euler.go:16-47) from the field held in brick 0, with respect to that initial
 * field and to the coefficient of every explicit (RHS.Terms) and implicit
 * (RHS.Denum) term of equation 0.  The transposed step is assembled term by term:
 * every map x -> IFFT(h FFT x)/N with h real and even is symmetric, so with
 * s = sigma / (1 - dt D) and Lam = FFT(lam_{n+1})
 *   lam_n   = Re IFFT(s Lam)/N + sum_t dt coef_t Re IFFT(s L^p_t Lam)/N * d b_t / d c
 *   dJ/dcoef_t (explicit) += Re sum_k conj(Lam) s dt L^p_t FFT(b_t) / N
 *   dJ/dcoef_t (implicit) += Re sum_k conj(Lam) dt L^p_t FFT(c_{n+1}) / (1 - dt D) / N
 * Supported: one field, Euler, concrete terms with real coefficients whose bricks
 * are the field or derived fields of it, implicit terms without a brick.  Every
 * state is kept (small grids only).  Leaves c_T in brick 0.  Returns 0, or -1 for
 * an unsupported model.
 */
int or_euler_adjoint(void *h, int nsteps, const double *target, double *loss, double *grad_c0,
                     double *grad_terms, double *grad_denum)
{
    or_model *m = (or_model *)h;
    long n = m->n;
    double dt = m->dt, f[3];
    if (m->stepper != 0 || m->nfields != 1 || m->neq < 1 || nsteps < 1) return -1;
    int t0 = m->eq_t_start[0], t1 = m->eq_t_start[1];
    int d0 = m->eq_d_start[0], d1 = m->eq_d_start[1];
    for (int t = t0; t < t1; ++t)
        if (m->terms[t].kind != TERM_CONCRETE || m->terms[t].coef_im != 0.0) return -1;
    for (int t = d0; t < d1; ++t)
        if (m->denum[t].kind != TERM_CONCRETE || m->denum[t].brick >= 0 || m->denum[t].coef_im != 0.0) return -1;
    for (int d = 0; d < m->nderived; ++d)
        for (int j = m->dfac_start[m->dterm_start[d]]; j < m->dfac_start[m->dterm_start[d + 1]]; ++j)
            if (m->dfac_field[j] != 0) return -1;
    int nb = m->nfields + m->nderived;
    double *states = (double *)malloc(sizeof(double) * n * (size_t)(nsteps + 1));
    for (int s = 0; s < nsteps; ++s) {
        for (long i = 0; i < n; ++i) states[(size_t)s * n + i] = creal(m->bricks[0][i]);
        euler_step(m);
    }
    for (long i = 0; i < n; ++i) states[(size_t)nsteps * n + i] = creal(m->bricks[0][i]);
    /* loss and lam_T */
    double *lam = (double *)malloc(sizeof(double) * n);
    double js = 0.0, jc = 0.0;
    for (long i = 0; i < n; ++i) {
        double r = states[(size_t)nsteps * n + i] - target[i];
        double v = r * r, t = js + v;
        if (fabs(js) >= fabs(v)) jc += (js - t) + v;
        else jc += (v - t) + js;
        js = t;
        lam[i] = 2.0 * r;
    }
    *loss = js + jc;
    for (int t = t0; t < t1; ++t) grad_terms[t - t0] = 0.0;
    for (int t = d0; t < d1; ++t) grad_denum[t - d0] = 0.0;
    zc *L = (zc *)malloc(sizeof(zc) * n), *C1 = (zc *)malloc(sizeof(zc) * n);
    zc *den = (zc *)malloc(sizeof(zc) * n), *val = (zc *)malloc(sizeof(zc) * n);
    zc **mu = (zc **)malloc(sizeof(zc *) * nb);
    double *sk = (double *)malloc(sizeof(double) * n), *lnew = (double *)malloc(sizeof(double) * n);
    for (int b = 0; b < nb; ++b) mu[b] = (zc *)malloc(sizeof(zc) * n);
    for (int s = nsteps - 1; s >= 0; --s) {
        const double *cs = states + (size_t)s * n, *cs1 = states + (size_t)(s + 1) * n;
        for (long i = 0; i < n; ++i) m->bricks[0][i] = cs[i];
        sync_derived(m);
        for (int b = 0; b < nb; ++b) fftnd_exec(m->ft, m->bricks[b], -1);
        get_denum(m, 0, den);
        for (long i = 0; i < n; ++i) {
            double sv = 1.0 / (1.0 - dt * creal(den[i]));
            if (m->filter_order > 0) {
                ref_freq(m->ndim, m->dims, i, f);
                double frad = sqrt(f[0] * f[0] + f[1] * f[1] + (m->ndim > 2 ? f[2] * f[2] : 0.0));
                sv *= or_vandeven_eval(m->vand, frad * 2.0 / OR_PI);
            }
            sk[i] = sv;
            L[i] = lam[i];
            C1[i] = cs1[i];
        }
        fftnd_exec(m->ft, L, -1);
        fftnd_exec(m->ft, C1, -1);
        for (int b = 0; b < nb; ++b)
            for (long i = 0; i < n; ++i) mu[b][i] = 0.0;
        for (long i = 0; i < n; ++i) mu[0][i] = sk[i] * L[i];
        for (int t = t0; t < t1; ++t) {
            or_term u = m->terms[t];
            u.coef_re = 1.0;
            u.coef_im = 0.0;
            eval_term(m, &u, val); /* L^p FFT(b_t), or L^p alone for a constant term */
            double g = 0.0;
            for (long i = 0; i < n; ++i) g += creal(conj(L[i]) * sk[i] * dt * val[i]);
            grad_terms[t - t0] += g / (double)n;
            if (m->terms[t].brick < 0) continue;
            u.brick = -1;
            eval_term(m, &u, val); /* L^p */
            zc *mb = mu[m->terms[t].brick];
            for (long i = 0; i < n; ++i) mb[i] += dt * m->terms[t].coef_re * sk[i] * creal(val[i]) * L[i];
        }
        for (int t = d0; t < d1; ++t) {
            or_term u = m->denum[t];
            u.coef_re = 1.0;
            u.coef_im = 0.0;
            eval_term(m, &u, val); /* L^p */
            double g = 0.0;
            for (long i = 0; i < n; ++i)
                g += creal(conj(L[i]) * dt * creal(val[i]) * C1[i]) / (1.0 - dt * creal(den[i]));
            grad_denum[t - d0] += g / (double)n;
        }
        fftnd_exec(m->ft, mu[0], +1);
        for (long i = 0; i < n; ++i) lnew[i] = creal(mu[0][i]) / (double)n;
        for (int d = 0; d < m->nderived; ++d) {
            zc *mb = mu[m->nfields + d];
            fftnd_exec(m->ft, mb, +1);
            for (long i = 0; i < n; ++i) lnew[i] += creal(mb[i]) / (double)n * derived_deriv(m, d, cs[i]);
        }
        memcpy(lam, lnew, sizeof(double) * n);
    }
    memcpy(grad_c0, lam, sizeof(double) * n);
    for (long i = 0; i < n; ++i) m->bricks[0][i] = states[(size_t)nsteps * n + i];
    for (int b = 0; b < nb; ++b) free(mu[b]);
    free(mu); free(sk); free(lnew); free(L); free(C1); free(den); free(val); free(lam); free(states);
    return 0;
}
