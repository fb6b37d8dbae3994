// engine.hip -- device plans of gopf's spectral time step and the C ABI (include/gpf.h).
//
// Two stepping plans compile from a parsed pf.Model:
//
//  * fused single-field path (Euler, one field, explicit terms that are
//    LAP^p of polynomials of the field, implicit LAP^p terms, optional
//    Vandeven filter) -- the Cahn-Hilliard / Allen-Cahn class.  One step is
//      2D: rows_fused (x) -> pencil_update (y)                          2 HBM passes
//      3D: rows_fused (x) -> pencils fwd (y) -> pencil_update (z)
//          -> pencils inv (y)                                           4 HBM passes
//    The state between steps is the x-half-spectrum of the field scaled by
//    1/nx, so the x-inverse of one step and the x-forward of the next share
//    one pass.
//
//  * general path (any number of fields, Euler, RK4 or ImplicitEuler,
//    ConcreteTerms on fields / derived fields / constants, SquaredGradient,
//    HomogeneousModulusLinElast, Vandeven filter): derived fields in real
//    space, R2C / C2R transforms, and one k-space update kernel per field in
//    the reference's sequential order (euler.go:27-39, rk4.go:29-127).
//    ImplicitEuler solves its step's residual equation with a Newton-Krylov
//    iteration whose vectors stay on the device (implicitEuler.go:68-208).
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/gpf.h"
#include "batch.h"
#include "common.h"
#include "fft_core.h"
#include "io.h"
#include "reduce.h"
#include "krylov.h"
#include "model.h"
#include "pointwise.h"
#include "spectral.h"

namespace gpf {


static thread_local std::string g_last_error;

template <typename T>
class Buf {
public:
    Buf() = default;
    explicit Buf(size_t count) { alloc(count); }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    Buf(Buf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    Buf& operator=(Buf&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    ~Buf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&p_), count * sizeof(T)));
        n_ = count;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    void swap(Buf& o) { std::swap(p_, o.p_); std::swap(n_, o.n_); }

private:
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* p_ = nullptr;
    size_t n_ = 0;
};

// exp(-2 pi i q / n), q in [0, n), computed in extended precision
template <typename T>
static std::vector<cpx<T>> twiddle_host(int n) {
    std::vector<cpx<T>> w(n);
    for (int q = 0; q < n; ++q) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)q / (long double)n;
        w[q].x = (T)cosl(a);
        w[q].y = (T)sinl(a);
    }
    return w;
}

// Bluestein tables of a length-n transform of direction dir (+-1) with convolution length m:
// the chirp w_j = exp(dir pi i j^2 / n), j < n (the exponent reduced mod 2n exactly, in
// integers, before the angle is formed), and the forward m-point transform of the
// zero-padded conj(w) (index j and m - j), divided by m.  Built in long double with an
// iterative radix-2 transform.
static int bluestein_length(int n) {
    int m = 1;
    while (m < 2 * n - 1) m *= 2;
    return m;
}

template <typename T>
static void bluestein_tables(int n, int dir, int m, std::vector<cpx<T>>& chirp, std::vector<cpx<T>>& filt) {
    const long double pi = 3.14159265358979323846264338327950288L;
    std::vector<std::complex<long double>> w(n), b(m, std::complex<long double>(0.0L, 0.0L));
    for (long j = 0; j < n; ++j) {
        const long q = (j * j) % (2L * n);
        const long double ang = (long double)dir * pi * (long double)q / (long double)n;
        w[j] = std::complex<long double>(cosl(ang), sinl(ang));
    }
    for (int j = 0; j < n; ++j) {
        b[j] = std::conj(w[j]);
        if (j > 0) b[m - j] = std::conj(w[j]);
    }
    // forward transform of b (sign -1), bit-reversal then butterflies
    for (int i = 1, r = 0; i < m; ++i) {
        int bit = m >> 1;
        for (; r & bit; bit >>= 1) r ^= bit;
        r ^= bit;
        if (i < r) std::swap(b[i], b[r]);
    }
    for (int len = 2; len <= m; len <<= 1) {
        const long double ang = -2.0L * pi / (long double)len;
        for (int i = 0; i < m; i += len)
            for (int k = 0; k < len / 2; ++k) {
                const std::complex<long double> tw(cosl(ang * k), sinl(ang * k));
                const std::complex<long double> u = b[i + k], v = b[i + k + len / 2] * tw;
                b[i + k] = u + v;
                b[i + k + len / 2] = u - v;
            }
    }
    chirp.resize(n);
    filt.resize(m);
    for (int j = 0; j < n; ++j) {
        chirp[j].x = (T)w[j].real();
        chirp[j].y = (T)w[j].imag();
    }
    for (int j = 0; j < m; ++j) {
        filt[j].x = (T)(b[j].real() / (long double)m);
        filt[j].y = (T)(b[j].imag() / (long double)m);
    }
}

// Vandeven table (pf/vandeven.go:13-27)
static void vandeven_table(int order, double* data) {
    data[0] = 1.0;
    const double prefactor = std::tgamma((double)(2 * order)) / (std::tgamma((double)order) * std::tgamma((double)order));
    const double dx = 1.0 / 999.0;
    for (int i = 1; i < 1000; ++i) {
        const double x = (double)i * dx;
        const double i2 = std::pow(x * (1 - x), (double)(order - 1));
        const double x1 = x - dx;
        const double i1 = std::pow(x1 * (1.0 - x1), (double)(order - 1));
        data[i] = data[i - 1] - prefactor * 0.5 * (i1 + i2) * dx;
    }
}

static double energy_density(const double* C, const double* eps) {
    double r = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) r += C[i * 27 + j * 9 + k * 3 + l] * eps[i * 3 + j] * eps[k * 3 + l];
    return 0.5 * r;
}

static bool real_coef(std::complex<double> c, double* out) {
    *out = c.real();
    return std::fabs(c.imag()) <= 1e-12 * std::max(1.0, std::fabs(c.real()));
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
struct Geometry {
    int ndim = 2;
    int dims[3] = {1, 1, 1};
    int nx = 1, ny = 1, nz = 1, nxh = 1;
    long n = 1, nmain = 0, nnyq = 0, nk = 0, rows = 0;
    KGrid kg;

    void init(const int* d, int nd) {
        if (nd != 2 && nd != 3) throw Error(GPF_EINVAL, "domain must be 2D or 3D");
        ndim = nd;
        for (int a = 0; a < nd; ++a) {
            if (d[a] < 2) throw Error(GPF_EINVAL, "domain extents must be >= 2");
            dims[a] = d[a];
        }
        if (nd == 2) {
            ny = d[0];
            nx = d[1];
            nz = 1;
        } else {
            // pfutil/fftWrap.go:42-74 indexes Freq with D1 as the fastest extent while FFTW
            // plans n = [D0, D1, D2]; the two agree only for cubic domains.
            if (d[0] != d[1] || d[1] != d[2])
                throw Error(GPF_EINVAL, "3D domains must be cubic (the reference's Freq convention, fftWrap.go:42-74, "
                                        "is only consistent for D0 = D1 = D2)");
            nz = d[0];
            ny = d[1];
            nx = d[2];
        }
        // the half spectrum keeps columns kx = 0 .. nx / 2; for odd nx the last one is an
        // ordinary frequency (its mirror is -kx), handled by the direct-DFT row kernels
        if (nx % 2 && nx > 64)
            throw Error(GPF_EINVAL, "an odd fastest extent must be at most 64 (no compiled odd-length schedule)");
        for (int v : {nx, ny, nz})
            if (v > 1 && !fft_size_supported(v))
                throw Error(GPF_EINVAL, "unsupported transform length " + std::to_string(v));
        nxh = nx / 2;
        n = (long)nx * ny * nz;
        rows = (long)ny * nz;
        nmain = rows * nxh;
        nnyq = rows;
        nk = nmain + nnyq;
        if (nk >= (1l << 32)) throw Error(GPF_EINVAL, "grid too large (the half spectrum must hold fewer than 2^32 values)");
        kg.nx = nx;
        kg.ny = ny;
        kg.nz = nz;
        kg.ndim = nd;
        kg.dxh = FastDiv::make((unsigned)nxh);
        kg.dy = FastDiv::make((unsigned)ny);
    }
    // KGrid::ftab on the host: float64(k) / float64(n), less 1 above 0.5 (fftWrap.go:57-74)
    std::vector<double> freq_table() const {
        auto wrap = [](int k, int m) {
            const double f = (double)k / (double)m;
            return (f > 0.5) ? f - 1.0 : f;
        };
        std::vector<double> t;
        for (int k = 0; k < ny; ++k) t.push_back(wrap(k, ny));
        for (int k = 0; k <= nxh; ++k) t.push_back(wrap(k, nx));
        for (int k = 0; k < nz; ++k) t.push_back(ndim == 3 ? wrap(k, nz) : 0.0);
        return t;
    }
};

// a block of tw lines fits: threads, and LDS for the exchange lines, the row pass's copy of
// the twiddle table and the monitor's FP64 partial sums (160 KiB per workgroup on gfx950)
static bool tw_fits(int n, int tw, size_t elem, int variant = 0) {
    const int T = fft_threads_per_line(n, variant);
    const size_t lds = ((size_t)(n + n / 16 + 1) * (tw + 1) + n) * elem + (size_t)kMaxBlockThreads * sizeof(double);
    return tw * T <= kMaxBlockThreads && lds <= 160 * 1024;
}

static int choose_tw(int n, int narr, size_t elem, int min_tw, int variant = 0, int threads = 256) {
    const int T = fft_threads_per_line(n, variant);
    int tw = 1;
    while (tw * 2 * T <= threads) tw *= 2;  // lines per block: a power of two
    if (tw > 64) tw = 64;
    while (tw < min_tw && tw * 2 * T <= 256) tw *= 2;
    while (tw * T > kMaxBlockThreads && tw > 1) tw /= 2;
    (void)narr;
    while (tw > 1 && (size_t)(n + n / 16 + 1) * (tw + 1) * elem > 80 * 1024) tw /= 2;
    return tw;
}

// lines per block of a general-path pass over `lines` lines: a small grid's pass is
// latency-bound, so fewer lines per block spread it over at least kSpreadBlocks workgroups
// (at 128^2 the row transform ran on 4 workgroups); large grids keep tw
constexpr long kSpreadBlocks = 256;
static int spread_tw(int tw, long lines, int min_tw = 1) {
    while (tw > min_tw && (lines + tw - 1) / tw < kSpreadBlocks) tw /= 2;
    return tw;
}

static Seg seg_y_main(const Geometry& g) {
    Seg s;
    s.na = g.nz; s.nb = g.nxh; s.sa = (long)g.ny * g.nxh; s.sb = 1; s.es = g.nxh;
    s.ksrc[0] = 1; s.ksrc[1] = 2; s.ksrc[2] = 0;
    return s;
}
static Seg seg_y_nyq(const Geometry& g) {
    Seg s;
    s.na = 1; s.nb = g.nz; s.sa = 0; s.sb = g.ny; s.es = 1;
    s.ksrc[0] = 3; s.kfix[0] = g.nxh; s.ksrc[1] = 2; s.ksrc[2] = 1;
    return s;
}
static Seg seg_z_main(const Geometry& g) {
    Seg s;
    s.na = g.ny; s.nb = g.nxh; s.sa = g.nxh; s.sb = 1; s.es = (long)g.ny * g.nxh;
    s.ksrc[0] = 1; s.ksrc[1] = 0; s.ksrc[2] = 2;
    return s;
}
static Seg seg_z_nyq(const Geometry& g) {
    Seg s;
    s.na = 1; s.nb = g.ny; s.sa = 0; s.sb = 1; s.es = g.ny;
    s.ksrc[0] = 3; s.kfix[0] = g.nxh; s.ksrc[1] = 1; s.ksrc[2] = 2;
    return s;
}

// ---------------------------------------------------------------------------
// compiled equation of one field (general path)
// ---------------------------------------------------------------------------
struct TermRef {
    int kind = kConcrete;
    double coef = 1.0;
    int pw = 0;
    int brick = -1;  // concrete: brick index
    int user = -1;
};

struct FieldEq {
    std::vector<TermRef> terms;
    KPoly denom;
    SvPoly sv;
    HessPoly hs;
};

// fused single-field plan
struct FusedPlan {
    bool ok = false;
    std::string why;
    std::vector<double> gpoly;
    int src_lap = 0;
    bool has_src = false;
    KPoly denom, lin;
};

static void kpoly_add(KPoly& p, double c, int pw) {
    for (int t = 0; t < p.n; ++t)
        if (p.pow[t] == pw) { p.coef[t] += c; return; }
    if (p.n >= kMaxKTerms) throw Error(GPF_EINVAL, "too many distinct LAP powers in one equation");
    p.coef[p.n] = c;
    p.pow[p.n] = pw;
    ++p.n;
}

// coef(t, &c) supplies each concrete term's real coefficient: the term's own value, or
// its derivative with respect to one scalar (the adjoint's parameter tables)
using CoefFn = std::function<bool(const TermDesc&, double*)>;

static FusedPlan compile_fused(const Model& m, int stepper,
                               const CoefFn& coef = [](const TermDesc& t, double* c) { return real_coef(t.coef, c); }) {
    FusedPlan fp;
    auto fail = [&](const std::string& w) { fp.ok = false; fp.why = w; return fp; };
    if (m.field_names.size() != 1) return fail("more than one field");
    if (stepper != GPF_EULER) return fail("stepper is not Euler");
    const auto& terms = m.rhs[0].first;
    const auto& denum = m.rhs[0].second;
    std::vector<double> g(kMaxPoly, 0.0);
    int src_lap = -1;
    for (const auto& t : terms) {
        if (t.kind != kConcrete) return fail("registered term");
        double c;
        if (!coef(t, &c)) return fail("complex coefficient");
        const int pw = t.has_lap ? t.lap : 0;
        if (pw < 0) return fail("negative LAP power");
        if (t.brick < 0) return fail("constant term");
        if (t.brick == 0) {
            kpoly_add(fp.lin, c, pw);
            continue;
        }
        const DerivedDesc& d = m.derived[t.brick - 1];
        if (d.kind == 2) return fail("noise source");
        if (src_lap >= 0 && src_lap != pw) return fail("explicit sources with different LAP powers");
        src_lap = pw;
        for (const auto& dt : d.terms) {
            double power = 0.0;
            for (const auto& f : dt.second) {
                if (f.first != 0) return fail("derived field of another field");
                power += f.second;
            }
            if (power != std::floor(power) || power < 0 || power >= kMaxPoly) return fail("non-integer power");
            g[(int)power] += c * dt.first;
        }
    }
    for (const auto& t : denum) {
        if (t.kind != kConcrete || t.brick >= 0) return fail("field-dependent implicit term");
        double c;
        if (!coef(t, &c)) return fail("complex coefficient");
        kpoly_add(fp.denom, c, t.has_lap ? t.lap : 0);
    }
    fp.has_src = src_lap >= 0;
    fp.src_lap = fp.has_src ? src_lap : 0;
    int deg = 0;
    for (int q = 0; q < kMaxPoly; ++q)
        if (g[q] != 0.0) deg = q + 1;
    fp.gpoly.assign(g.begin(), g.begin() + deg);
    fp.ok = true;
    return fp;
}

bool fused_shape(const Model& m, FusedShape* out, std::string* why) {
    FusedPlan fp = compile_fused(m, GPF_EULER);
    if (!fp.ok) {
        if (why) *why = fp.why;
        return false;
    }
    out->gpoly = fp.gpoly;
    out->src_lap = fp.src_lap;
    out->has_src = fp.has_src;
    out->denom = fp.denom;
    out->lin = fp.lin;
    return true;
}

bool same_structure(const FusedShape& a, const FusedShape& b) {
    if (a.has_src != b.has_src || a.src_lap != b.src_lap || a.denom.n != b.denom.n || a.lin.n != b.lin.n) return false;
    for (int t = 0; t < a.denom.n; ++t)
        if (a.denom.pow[t] != b.denom.pow[t]) return false;
    for (int t = 0; t < a.lin.n; ++t)
        if (a.lin.pow[t] != b.lin.pow[t]) return false;
    return true;
}

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
class EngineBase {
public:
    virtual ~EngineBase() = default;
    virtual void step() = 0;
    virtual void steps(long n) = 0;  // n steps; the fused path replays captured graphs
    // everything the first steps would set up, without advancing the state: the
    // autotuner, the monitor buffers for `upcoming` steps and the replayed step graph;
    // returns the host milliseconds it took
    virtual double prepare(long upcoming) = 0;
    virtual void on_step_finished() = 0;
    virtual void get_field(int idx, double* out) = 0;
    virtual void set_field(int idx, const double* in) = 0;
    virtual void energy(int idx, int npoly, const double* poly, double kappa, double* e, double* mass) = 0;
    virtual void monitor(int npoly, const double* poly, double kappa) = 0;
    // FP64 sum, sum of squares, min, max and non-finite count of field idx (reduce.h order)
    virtual void reduce(int idx, double* out) = 0;
    virtual long series(double* e, double* m, long cap) = 0;
    virtual bool fused() const = 0;
    virtual void enable_timing(bool on) = 0;
    virtual int pass_times(int maxpass, double* ms, long* launches, char* names, int namelen) = 0;
    virtual hipStream_t stream() const = 0;
    virtual long size() const = 0;
    virtual int tuning(int* tws, int maxn) = 0;
    virtual void set_tuning(const int* tws, int n) = 0;
    virtual void set_filter(const double* table) = 0;
    virtual void adjoint(long nsteps, long every, const double* target, double* loss, double* grad_field,
                         int nscal, const char* const* names, double* grad_scalars) = 0;
    virtual void set_newton_krylov(int maxiter, double step_size, double tol, int stencil, int restart) = 0;
    virtual void newton_stats(int* newton_iters, int* krylov_iters, int* converged, double* residual) const = 0;
    virtual void save_fields(const char* prefix, int epoch, int format) = 0;
    virtual void sdd_init(const double* init, const double* final, long n) = 0;
    virtual void sdd_set_orientation(const double* v, long n) = 0;
    virtual void sdd_set_params(double alpha, double tau_o, double tau_l, double init_len, double min_len) = 0;
    virtual void sdd_monitor(double* out) = 0;
    virtual void sdd_orientation(double* out, long n) = 0;
    virtual double sdd_dimer_length(double t) const = 0;
    virtual void flush_io() = 0;
    virtual void save_csv(const char* prefix, int epoch) = 0;
    // the persistent 2D step (k_fused_persist): mode -1 where it is measured faster, 0 never,
    // 1 wherever admissible; threads per block 0 (64) or 64, 128, 256, 512
    virtual void set_persistent(int mode, int threads) = 0;
    virtual long persistent_steps() const = 0;
    // errors a finished asynchronous launch left behind (called after a stream synchronisation)
    virtual void after_sync() = 0;
};

template <typename T>
class Engine : public EngineBase {
public:
    Engine(const Model& m, const int* dims, int ndim, double dt, const gpf_options& o);
    ~Engine() override;

    void step() override;
    void steps(long n) override;
    double prepare(long upcoming) override;
    void on_step_finished() override;
    void get_field(int idx, double* out) override;
    void set_field(int idx, const double* in) override;
    void energy(int idx, int npoly, const double* poly, double kappa, double* e, double* mass) override;
    void monitor(int npoly, const double* poly, double kappa) override;
    void reduce(int idx, double* out) override;
    long series(double* e, double* m, long cap) override;
    bool fused() const override { return fp_.ok; }
    void enable_timing(bool on) override { timing_ = on; }
    int pass_times(int maxpass, double* ms, long* launches, char* names, int namelen) override;
    hipStream_t stream() const override { return st_; }
    long size() const override { return g_.n; }
    int tuning(int* tws, int maxn) override;
    void set_tuning(const int* tws, int n) override;
    void set_filter(const double* table) override;
    void adjoint(long nsteps, long every, const double* target, double* loss, double* grad_field, int nscal,
                 const char* const* names, double* grad_scalars) override;
    void set_newton_krylov(int maxiter, double step_size, double tol, int stencil, int restart) override;
    void newton_stats(int* newton_iters, int* krylov_iters, int* converged, double* residual) const override;
    void save_fields(const char* prefix, int epoch, int format) override;
    void sdd_init(const double* init, const double* final, long n) override;
    void sdd_set_orientation(const double* v, long n) override;
    void sdd_set_params(double alpha, double tau_o, double tau_l, double init_len, double min_len) override;
    void sdd_monitor(double* out) override;
    void sdd_orientation(double* out, long n) override;
    double sdd_dimer_length(double t) const override;
    void flush_io() override;
    void save_csv(const char* prefix, int epoch) override;
    void set_persistent(int mode, int threads) override;
    long persistent_steps() const override { return persist_steps_; }
    void after_sync() override { check_persist(); }

private:
    // transforms
    const cpx<T>* tw(int n);
    // stage (optional, inverse passes): the fields' RK4 stage update done as they are loaded
    // (launch_pencils_stage); src is then unused
    void pencils(const std::vector<cpx<T>*>& dst, const std::vector<const cpx<T>*>& src, int axis, int dir,
                 const PencilStageArgs<T>* stage = nullptr);
    // prog (optional): per input, a dense derived-field program evaluated while loading
    // instead of reading the input array
    void r2c(const std::vector<const T*>& in, const std::vector<cpx<T>*>& out, T in_scale,
             const std::vector<const DerivedArgs<T>*>* prog = nullptr);
    void c2r(const std::vector<const cpx<T>*>& in, const std::vector<T*>& out, T scale,
             const double* add_mean = nullptr, const PencilStageArgs<T>* stage = nullptr);
    cpx<T>* kscratch(int i);
    T* rscratch(int i);
    // timing
    void tbegin(int pass);
    void tend(int pass);
    // fused path
    void fused_step();
    void prepare_fused(long upcoming);  // allocations ahead of launches (never during graph capture)
    // the launches of step pos of count steps enqueued back to back (capturable)
    void enqueue_fused(int pos = 0, int count = 1);
    void capture_graph();               // instantiates graph_: kGraphSteps steps of enqueue_fused
    void drop_graph();
    hipGraphExec_t graph_ = nullptr;
    bool graph_mon_ = false;  // the monitor setting the graph was captured with
    static constexpr int kGraphSteps = 10;
    // largest side taking the persistent step by default: faster than the replayed per-pass
    // graphs at 32^2-128^2 (FP64 128^2 with the energy series 12.9 against 17.0 us per step,
    // 32^2 10.7 against 14.6), slower at 256^2 (FP64 17.0 us with 128 threads per block
    // against 13.7-14.3 us for the tuned graphs)
    static constexpr int kPersistAutoMax = 128;
    // persistent 2D step: a run of steps in one cooperative launch (spectral_persist.h); its
    // plan (the passes' arguments) is rebuilt whenever the graph is dropped.  Modes: -1 where
    // measured faster, 0 never, 1 wherever admissible, 2 as 1 with the agent-fenced barrier
    // (the layout taken when the device does not place the participants on one XCD)
    bool persist_eligible();
    void build_persist();
    long run_persist(long n);  // returns the steps left to the per-pass launches
    void check_persist();
    int persist_mode_ = -1, persist_threads_ = 0;
    int coop_ = -1;  // the device's cooperative-launch attribute (-1 not yet read)
    int local_ = -1;  // participants on one XCD: -1 not yet seen, 1 seen, 0 not the device's placement
    long persist_steps_ = 0;
    bool persist_ok_ = false;
    long pnr_ = 0, pnu_ = 0;  // block counts of the plan's row and update passes
    Buf<unsigned> pbar_;
    Buf<double> ppa_bulk_, ppa_mass_, ppc_eg_;
    Buf<cpx<T>> ptab_;  // the update pass's multipliers of every k-point
    Buf<double> pm0_;   // m1(0), m2(0)
    int* pfail_ = nullptr;      // mapped host flag set by a failed device-wide barrier
    int* pfail_dev_ = nullptr;
    PersistLaunch<T> pl_;
    RowFusedArgs<T> rows_args(cpx<T>* W, double* mean, int tw, int var);
    PencilUpdateArgs<T> update_args(cpx<T>* S, cpx<T>* W, double* mean, int tw);
    void prime();
    void pencils_tw(const std::vector<cpx<T>*>& dst, int axis, int dir, int tw, int lpt = 1, int var = 0);
    void autotune();
    void warm(double ms_min);  // the step's passes on scratch copies for at least ms_min ms
    bool autotune_ = false, tuned_ = false, warmed_ = false;
    int tw_rows_ = 0, tw_pf_ = 0, tw_pu_ = 0, tw_pi_ = 0;
    int rows_var_ = 0;  // transform schedule of the row pass (0 plain, 1 low-register)
    int lpt_pf_ = 1, lpt_pi_ = 1;  // lines per thread of the forward / inverse pencil passes
    T* c_out_ = nullptr;           // when set, the next row pass also writes the real field c here
    int var_pf_ = 0, var_pi_ = 0;  // transform schedules of the forward / inverse pencil passes
    // 3D: axis of the fused forward-update-inverse pass (2 = z, 1 = y); the other
    // of y / z is transformed forward before and inverse after it
    int upd_axis_ = 2;
    int upd_axis() const { return (g_.ndim == 3) ? upd_axis_ : 1; }
    int other_axis() const { return upd_axis() == 2 ? 1 : 2; }
    int axis_len(int axis) const { return axis == 2 ? g_.nz : g_.ny; }
    // general path
    void general_step_euler();
    void general_step_rk4();
    void sync_from_spectra(const PencilStageArgs<T>* stage = nullptr);
    bool fused_stage_args(int stage, PencilStageArgs<T>* out);
    void ensure_k_fields();
    void sync_derived();
    void eval_user_terms_for(int eq);
    void eval_elastic();
    KElasticArgs<T> elastic_args(size_t e) const;
    void rhs_update(int eq, cpx<T>* dst, const cpx<T>* cur, double alpha, double beta, double gamma, bool filt,
                    const cpx<T>* extra);
    template <class A>
    void terms_into(int eq, A& a);
    // ImplicitEuler (FP64): the residual F(x) of the exponential-integrator step, its
    // finite-difference directional derivative and a restarted GMRES solve of J dx = -F
    void implicit_step();
    void implicit_residual(const double* x, double* out);
    void implicit_jv(const double* v, double* out);
    int implicit_gmres();
    double dev_scalar(const double* p);
    struct NewtonKrylov {
        int maxiter = 50;       // DefaultNonLinSolver (implicitEuler.go:221-229)
        double step = 1e-3;
        double tol = 1e-7;
        int stencil = 6;
        int restart = 30;       // GMRES restart length (device choice)
    } nk_;
    int nk_newton_ = 0, nk_krylov_ = 0, nk_conv_ = 0;
    double nk_res_ = 0.0;
    std::vector<Buf<cpx<T>>> ORIG_, RP_, RES_;
    // Newton-Krylov vectors of length nf n: iterate, residual, update, perturbed point,
    // residual there, Arnoldi work vector, GMRES basis (restart + 1 vectors)
    Buf<double> nx_, nfx_, ndx_, nxp_, nfp_, nw_, nv_, npart_, ndev_;
    std::vector<double> nhost_;
    // SDD (sdd.go:84-301): orientation (nf n, unit norm), its spectra, the right-hand sides at
    // the dimer's ends, the real torque, partial sums, device scalars and monitors
    void sdd_step();
    double* npart_sdd() { return sdd_part_.get() + 3 * kSddBlocks; }  // krylov_blocks(n) <= 1024 partials
    bool sdd_ready_ = false;
    double sdd_alpha_ = 0.5, sdd_tau_o_ = 1.0, sdd_tau_l_ = 1.0, sdd_len0_ = 0.0, sdd_min_len_ = 0.0;
    long sdd_step_ = 0;
    Buf<double> sdd_v_, sdd_tau_, sdd_part_, sdd_sc_, sdd_mon_;
    std::vector<Buf<cpx<T>>> sdd_ov_, sdd_rs_, sdd_re_;
    // field snapshots (Float64IO / Uint8IO)
    std::unique_ptr<SnapshotWriter> io_;
    Buf<double> io_part_, io_mm_;

    Geometry g_;
    hipStream_t st_ = nullptr;
    double dt_;
    int stepper_;
    int nf_ = 0, nd_ = 0;
    Model mcopy_;
    FusedPlan fp_;
    std::vector<FieldEq> eqs_;
    Buf<double> filter_;
    Buf<double> ftab_;  // KGrid::ftab
    bool row_sync_ = false;  // sync_from_spectra applies (dense derived fields of <= 2 fields)
    bool has_filter_ = false;
    std::vector<std::pair<int, Buf<cpx<T>>>> tw_;
    // general-path storage
    std::vector<Buf<T>> F_, Dr_;
    std::vector<Buf<cpx<T>>> K_, Tsq_, Tel_, INIT_, FINAL_, KF_;
    // RK4 stage factor tables per field: [3 i] gamma = dt / 2, [3 i + 1] gamma = dt,
    // [3 i + 2] gamma = dt with the filter (empty when the equation has a term that needs the
    // frequency vector); built once per plan, dropped when the filter changes
    std::vector<Buf<T>> stab_;
    bool stab_ok_ = false;
    void build_stage_tables();
    std::vector<Buf<T>> Ktab_;  // K(k) of each elastic term, nk values (built in FP64, stored in T)
    std::vector<Buf<T>> lag_;
    // VolumeConservingLP: indicator brick, field, device state {Multiplier, CurrentIntegral,
    // IsFirstUpdate} per occurrence, and the partial sums of its field integral
    std::vector<int> vc_brick_, vc_field_;
    std::vector<Buf<double>> vc_state_;
    Buf<double> vc_part_;
    std::vector<DerivedArgs<T>> dargs_;
    std::vector<unsigned> dcount_;  // syncs so far of each derived field (the noise counter)
    std::vector<std::vector<int>> cn_bricks_;  // ConservativeNoise: current-field bricks per occurrence
    bool kvalid_ = false;
    std::vector<Buf<cpx<T>>> ks_;
    std::vector<Buf<T>> rs_;
    // fused-path storage: the state spectrum S = FFT(c - mean) / N and one work array W,
    // which between steps holds S inverse-transformed along the update axis
    Buf<cpx<T>> S_, W_;
    Buf<double> mean_;          // FP32 fused path: FP64 mean of the field (the arrays hold c - mean)
    bool split_mean_ = false;
    double host_mean();
    // monitors
    bool mon_ = false;
    std::vector<double> mon_poly_;
    double mon_kappa_ = 0.0;
    Buf<double> pa_bulk_, pa_mass_, pc_eg_, series_, dpoly_;
    Buf<long> series_count_;  // device slot counter of the series (replayable by a graph)
    long nblk_a_ = 0, nblk_c_ = 0, series_cap_ = 0, series_n_ = 0;
    // timing
    bool timing_ = false;
    std::vector<std::string> pass_names_;
    std::vector<double> pass_ms_;
    std::vector<long> pass_cnt_;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending_;
    hipEvent_t cur_ev_ = nullptr;
};

template <typename T>
Engine<T>::Engine(const Model& m, const int* dims, int ndim, double dt, const gpf_options& o)
    : autotune_(o.autotune != 0), dt_(dt), stepper_(o.stepper), mcopy_(m) {
    g_.init(dims, ndim);
    HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    {
        const std::vector<double> ft = g_.freq_table();
        ftab_.alloc(ft.size());
        HIPCHK(hipMemcpy(ftab_.get(), ft.data(), ft.size() * sizeof(double), hipMemcpyHostToDevice));
        g_.kg.ftab = ftab_.get();
    }
    nf_ = (int)m.field_names.size();
    nd_ = (int)m.derived.size();
    if (nf_ == 0) throw Error(GPF_EINVAL, "Model: No fields added");
    if (nf_ > kMaxFields) throw Error(GPF_EINVAL, "too many fields");
    if (m.rhs.size() != m.equations.size() || m.rhs.empty())
        throw Error(GPF_ESTATE, "model is not initialised (call gpf_model_init)");
    if ((int)m.rhs.size() < nf_)
        throw Error(GPF_EINVAL, "every field needs an equation (RHS[i] updates Fields[i], euler.go:27-39)");
    for (int i = 0; i < nf_; ++i)
        if ((long)m.field_data[i].size() != g_.n)
            throw Error(GPF_EINVAL, "solver: Inconsistent domain size and number of grid points");
    if (o.filter_order > 0) {
        std::vector<double> tab(1000);
        vandeven_table(o.filter_order, tab.data());
        set_filter(tab.data());
    }
    if (o.fused) fp_ = compile_fused(m, stepper_);
    if (fp_.ok) {
        S_.alloc(g_.nk);
        W_.alloc(g_.nk);
        mean_.alloc(1);
        split_mean_ = (sizeof(T) == 4);
        pass_names_ = {"rows_fused", "pencils_fwd", "pencil_update", "pencils_inv"};
    } else {
        // compile every equation for the general path
        for (int i = 0; i < nf_; ++i) {
            FieldEq fe;
            for (const auto& t : m.rhs[i].first) {
                TermRef r;
                r.kind = t.kind;
                if (!real_coef(t.coef, &r.coef)) throw Error(GPF_EINVAL, "complex scalar coefficients are not supported (real fields)");
                r.pw = (t.kind == kConcrete) ? (t.has_lap ? t.lap : 0) : t.lap;
                if (r.pw < 0) throw Error(GPF_EINVAL, "negative LAP power");
                r.brick = t.brick;
                r.user = t.user;
                fe.terms.push_back(r);
            }
            if (fe.terms.size() > (size_t)kMaxUTerms) throw Error(GPF_EINVAL, "too many terms in one equation");
            for (const auto& t : m.rhs[i].second) {
                if (t.kind == kSpectralViscosity) {
                    const UserTerm& u = m.user_table[kSpectralViscosity][t.user];
                    if (fe.sv.n >= kMaxSv) throw Error(GPF_EINVAL, "too many SpectralViscosity terms in one equation");
                    if (u.power < 0) throw Error(GPF_EINVAL, "SpectralViscosity.Power must be non-negative");
                    const int q = fe.sv.n++;
                    fe.sv.coef[q] = t.coef.real() * -u.eps;
                    fe.sv.thr[q] = u.threshold;
                    fe.sv.pw[q] = u.power;
                    fe.sv.lap[q] = t.lap;
                    continue;
                }
                if (t.kind == kTensorialHessian) {
                    const UserTerm& u = m.user_table[kTensorialHessian][t.user];
                    if (fe.hs.n >= kMaxHess) throw Error(GPF_EINVAL, "too many TensorialHessian terms in one equation");
                    const int d = (u.hess_len == 9) ? 3 : 2;
                    if (u.hess_len != d * d || d < g_.ndim)
                        throw Error(GPF_EINVAL, "TensorialHessian.K must hold ndim x ndim coefficients");
                    const int q = fe.hs.n++;
                    fe.hs.coef[q] = t.coef.real();
                    fe.hs.lap[q] = t.lap;
                    fe.hs.kd[q] = d;
                    for (int k = 0; k < 9; ++k) fe.hs.K[q][k] = u.hess[k];
                    continue;
                }
                if (t.kind != kConcrete || t.brick >= 0)
                    throw Error(GPF_EINVAL, "implicit terms that depend on fields cannot be evaluated on the device "
                                            "(only scalar * LAP^p terms are implicit)");
                double c;
                if (!real_coef(t.coef, &c)) throw Error(GPF_EINVAL, "complex scalar coefficients are not supported");
                kpoly_add(fe.denom, c, t.has_lap ? t.lap : 0);
            }
            eqs_.push_back(fe);
        }
        for (int i = 0; i < nf_; ++i) F_.emplace_back(g_.n);
        for (int d = 0; d < nd_; ++d) Dr_.emplace_back(g_.n);
        for (int b = 0; b < nf_ + nd_; ++b) K_.emplace_back(g_.nk);
        for (size_t s = 0; s < m.user_table[kSquaredGradient].size(); ++s) Tsq_.emplace_back(g_.nk);
        for (size_t e = 0; e < m.user_table[kHomogeneousModulus].size(); ++e) {
            Tel_.emplace_back(g_.nk);
            Ktab_.emplace_back(g_.nk);
            HIPCHK(launch_kelastic_table<T>(elastic_args(e), Ktab_.back().get(), st_));
            lag_.emplace_back(g_.n);
            HIPCHK(hipMemsetAsync(lag_.back().get(), 0, g_.n * sizeof(T), st_));
            bool found = false;
            for (int f = 0; f < nf_; ++f) found |= (m.field_names[f] == m.user_table[kHomogeneousModulus][e].field);
            if (!found) throw Error(GPF_EINVAL, "HomogeneousModulusLinElast field is not a field");
        }
        for (const auto& s : m.user_table[kSquaredGradient]) {
            if (!m.is_field_name(s.field)) throw Error(GPF_EINVAL, "SquaredGradient field is not a field");
        }
        for (const auto& v : m.user_table[kVolumeConserving]) {
            if (!m.is_field_name(v.field)) throw Error(GPF_EINVAL, "VolumeConservingLP field is not a field");
            if (v.num_nodes < 0 || v.num_nodes > g_.n)
                throw Error(GPF_EINVAL, "VolumeConservingLP.NumNodes exceeds the number of nodes");
            vc_brick_.push_back(m.brick_index(v.indicator));
            vc_field_.push_back(m.brick_index(v.field));
            vc_state_.emplace_back(3);
            const double init[3] = {0.0, 0.0, 1.0};  // Multiplier, CurrentIntegral, IsFirstUpdate
            HIPCHK(hipMemcpy(vc_state_.back().get(), init, sizeof(init), hipMemcpyHostToDevice));
        }
        if (!m.user_table[kVolumeConserving].empty()) vc_part_.alloc(1024);
        for (const auto& c : m.user_table[kConservativeNoise]) {
            std::vector<int> b;
            for (int k = 0; k < c.cn_dim; ++k) b.push_back(m.brick_index(c.cn_names[k]));
            if (c.cn_dim > g_.ndim) throw Error(GPF_EINVAL, "ConservativeNoise.Dim exceeds the grid's dimension");
            cn_bricks_.push_back(b);
        }
        if (stepper_ == GPF_RK4)
            for (int i = 0; i < nf_; ++i) {
                INIT_.emplace_back(g_.nk);
                FINAL_.emplace_back(g_.nk);
                KF_.emplace_back(g_.nk);
            }
        if (stepper_ == GPF_SDD) {
            if (sizeof(T) != sizeof(double)) throw Error(GPF_EINVAL, "SDD runs in FP64 (GPF_FP64)");
            if (has_filter_) throw Error(GPF_EINVAL, "SDD: Does not support modal filters");
            for (int i = 0; i < nf_; ++i) {
                sdd_ov_.emplace_back(g_.nk);
                sdd_rs_.emplace_back(g_.nk);
                sdd_re_.emplace_back(g_.nk);
            }
            sdd_v_.alloc((size_t)nf_ * g_.n);
            sdd_tau_.alloc((size_t)nf_ * g_.n);
            sdd_part_.alloc(3 * kSddBlocks + 1024);
            sdd_sc_.alloc(16);
            sdd_mon_.alloc(8);
            HIPCHK(hipMemsetAsync(sdd_mon_.get(), 0, 8 * sizeof(double), st_));
        }
        if (stepper_ == GPF_IMPLICIT_EULER) {
            if (sizeof(T) != sizeof(double)) throw Error(GPF_EINVAL, "ImplicitEuler runs in FP64 (GPF_FP64)");
            for (int i = 0; i < nf_; ++i) {
                ORIG_.emplace_back(g_.nk);
                RP_.emplace_back(g_.nk);
                RES_.emplace_back(g_.nk);
            }
            const long M = (long)nf_ * g_.n;
            nx_.alloc(M);
            nfx_.alloc(M);
            ndx_.alloc(M);
            nxp_.alloc(M);
            nfp_.alloc(M);
            nw_.alloc(M);
            npart_.alloc(krylov_blocks(M) * (kMaxKrylov + 1));
            ndev_.alloc(4 * (kMaxKrylov + 1));
        }
        // derived-field programs
        for (int d = 0; d < nd_; ++d) {
            const DerivedDesc& dd = m.derived[d];
            DerivedArgs<T> a;
            a.nf = nf_;
            a.n = g_.n;
            a.out = Dr_[d].get();
            if ((int)dd.terms.size() > kMaxDTerms) throw Error(GPF_EINVAL, "derived field with too many terms");
            a.nterm = (int)dd.terms.size();
            for (int t = 0; t < a.nterm; ++t) {
                a.coef[t] = dd.terms[t].first;
                for (int f = 0; f < kMaxFields; ++f) a.pw[t][f] = 0.0;
                for (const auto& fac : dd.terms[t].second) a.pw[t][fac.first] += fac.second;
            }
            for (int f = 0; f < nf_; ++f) a.f[f] = F_[f].get();
            make_dense_derived(a);
            if (dd.kind == 2) {
                a.noise = 1;
                a.dense = 0;
                a.noise_std = dd.noise_std;
                a.seed = dd.noise_seed;
                a.stream = (unsigned)d;
            }
            dargs_.push_back(a);
            dcount_.push_back(0);
        }
        row_sync_ = nf_ <= 2 && nd_ >= 1 && nd_ <= kMaxArr;
        for (int d = 0; d < nd_; ++d) row_sync_ = row_sync_ && dargs_[d].dense && !dargs_[d].noise;
        pass_names_ = {"general_step"};
    }
    pass_ms_.assign(pass_names_.size(), 0.0);
    pass_cnt_.assign(pass_names_.size(), 0);
    for (int i = 0; i < nf_; ++i) {
        std::vector<double> buf(2 * g_.n);
        for (long j = 0; j < g_.n; ++j) {
            buf[2 * j] = m.field_data[i][j].real();
            buf[2 * j + 1] = m.field_data[i][j].imag();
        }
        set_field(i, buf.data());
    }
}

template <typename T>
Engine<T>::~Engine() {
    io_.reset();  // writes what is queued
    drop_graph();
    if (st_) {
        (void)hipStreamSynchronize(st_);
        for (auto& p : pending_) {
            (void)hipEventDestroy(p.second.first);
            (void)hipEventDestroy(p.second.second);
        }
        (void)hipStreamDestroy(st_);
    }
    if (pfail_) (void)hipHostFree(pfail_);
}

template <typename T>
const cpx<T>* Engine<T>::tw(int n) {
    for (auto& p : tw_)
        if (p.first == n) return p.second.get();
    auto h = twiddle_host<T>(n);
    Buf<cpx<T>> b(n);
    HIPCHK(hipMemcpy(b.get(), h.data(), n * sizeof(cpx<T>), hipMemcpyHostToDevice));
    tw_.emplace_back(n, std::move(b));
    return tw_.back().second.get();
}

template <typename T>
cpx<T>* Engine<T>::kscratch(int i) {
    while ((int)ks_.size() <= i) ks_.emplace_back(g_.nk);
    return ks_[i].get();
}

template <typename T>
T* Engine<T>::rscratch(int i) {
    while ((int)rs_.size() <= i) rs_.emplace_back(g_.n);
    return rs_[i].get();
}

template <typename T>
void Engine<T>::tbegin(int pass) {
    if (!timing_) return;
    hipEvent_t a;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventRecord(a, st_));
    cur_ev_ = a;
    (void)pass;
}

template <typename T>
void Engine<T>::tend(int pass) {
    if (!timing_) return;
    hipEvent_t b;
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(b, st_));
    pending_.push_back({pass, {cur_ev_, b}});
}

template <typename T>
int Engine<T>::pass_times(int maxpass, double* ms, long* launches, char* names, int namelen) {
    HIPCHK(hipStreamSynchronize(st_));
    for (auto& p : pending_) {
        float e = 0.f;
        HIPCHK(hipEventElapsedTime(&e, p.second.first, p.second.second));
        pass_ms_[p.first] += e;
        pass_cnt_[p.first] += 1;
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    pending_.clear();
    const int np = (int)pass_names_.size();
    for (int i = 0; i < np && i < maxpass; ++i) {
        if (ms) ms[i] = pass_ms_[i];
        if (launches) launches[i] = pass_cnt_[i];
        if (names && namelen > 0) {
            std::strncpy(names + (size_t)i * namelen, pass_names_[i].c_str(), namelen - 1);
            names[(size_t)i * namelen + namelen - 1] = 0;
        }
        pass_ms_[i] = 0.0;
        pass_cnt_[i] = 0;
    }
    return np;
}

template <typename T>
void Engine<T>::pencils(const std::vector<cpx<T>*>& dst, const std::vector<const cpx<T>*>& src, int axis, int dir,
                        const PencilStageArgs<T>* stage) {
    const int n = (axis == 1) ? g_.ny : g_.nz;
    if (stage && (dir != 1 || dst.size() > (size_t)kMaxArr))
        throw Error(GPF_EINVAL, "fused stage pencils: one inverse pass of at most 4 fields");
    for (size_t q0 = 0; q0 < dst.size(); q0 += kMaxArr) {
        PencilArgs<T> a;
        a.tw = tw(n);
        a.narr = (int)std::min(dst.size() - q0, (size_t)kMaxArr);
        a.seg[0] = (axis == 1) ? seg_y_main(g_) : seg_z_main(g_);
        a.seg[1] = (axis == 1) ? seg_y_nyq(g_) : seg_z_nyq(g_);
        a.nseg = 2;
        // about 128 threads per block: at 192^3 (C4) 8 lines per block beat 16 by 1 % (FP32)
        // and 3 % (FP64) per step
        a.TW = spread_tw(choose_tw(n, a.narr, sizeof(cpx<T>), 8, 0, 128),
                         (long)a.seg[0].na * a.seg[0].nb + (long)a.seg[1].na * a.seg[1].nb);
        for (int q = 0; q < a.narr; ++q) {
            a.arr[q][0] = dst[q0 + q];
            a.arr[q][1] = dst[q0 + q] + g_.nmain;
            if (!src.empty() && src[q0 + q]) {
                a.src[q][0] = src[q0 + q];
                a.src[q][1] = src[q0 + q] + g_.nmain;
            }
        }
        if (stage) {
            if constexpr (sizeof(T) == 8) {
                PencilStageArgs<T> sa = *stage;
                sa.p = a;
                HIPCHK(launch_pencils_stage<T>(sa, n, st_));
                continue;
            }
        }
        HIPCHK(launch_pencils<T>(a, n, dir, st_));
    }
}

template <typename T>
void Engine<T>::r2c(const std::vector<const T*>& in, const std::vector<cpx<T>*>& out, T in_scale,
                    const std::vector<const DerivedArgs<T>*>* prog) {
    for (size_t q0 = 0; q0 < in.size(); q0 += kMaxArr) {
        RowR2CArgs<T> a;
        a.tw = tw(g_.nx);
        a.narr = (int)std::min(in.size() - q0, (size_t)kMaxArr);
        for (int q = 0; q < a.narr; ++q) {
            a.in[q] = in[q0 + q];
            a.main[q] = out[q0 + q];
            a.nyq[q] = out[q0 + q] + g_.nmain;
            const DerivedArgs<T>* d = prog ? (*prog)[q0 + q] : nullptr;
            if (d) {
                a.poly[q] = 1;
                a.pf0[q] = d->f[0];
                a.pf1[q] = (d->nf > 1) ? d->f[1] : nullptr;
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j) a.dc[q][i][j] = d->dc[i][j];
            }
        }
        a.rows = g_.rows;
        a.nx = g_.nx;
        a.TW = spread_tw(choose_tw(g_.nx, (a.narr + 1) / 2, sizeof(cpx<T>), 1), g_.rows);
        a.in_scale = in_scale;
        HIPCHK(launch_rows_r2c<T>(a, st_));
    }
    pencils(out, {}, 1, -1);
    if (g_.ndim == 3) pencils(out, {}, 2, -1);
}

template <typename T>
void Engine<T>::c2r(const std::vector<const cpx<T>*>& in, const std::vector<T*>& out, T scale,
                    const double* add_mean, const PencilStageArgs<T>* stage) {
    if (stage && in.size() > (size_t)kMaxArr) throw Error(GPF_EINVAL, "fused stage pencils: at most 4 fields");
    for (size_t q0 = 0; q0 < in.size(); q0 += kMaxArr) {
        const int na = (int)std::min(in.size() - q0, (size_t)kMaxArr);
        std::vector<cpx<T>*> scr;
        std::vector<const cpx<T>*> src;
        for (int q = 0; q < na; ++q) {
            scr.push_back(kscratch(q));
            src.push_back(in[q0 + q]);
        }
        if (g_.ndim == 3) {
            pencils(scr, src, 2, +1, stage);
            pencils(scr, {}, 1, +1);
        } else {
            pencils(scr, src, 1, +1, stage);
        }
        RowC2RArgs<T> a;
        a.tw = tw(g_.nx);
        a.narr = na;
        for (int q = 0; q < na; ++q) {
            a.main[q] = scr[q];
            a.nyq[q] = scr[q] + g_.nmain;
            a.out[q] = out[q0 + q];
        }
        a.rows = g_.rows;
        a.nx = g_.nx;
        a.TW = spread_tw(choose_tw(g_.nx, (na + 1) / 2, sizeof(cpx<T>), 1), g_.rows);
        a.scale = scale;
        a.add_mean = (q0 == 0) ? add_mean : nullptr;
        HIPCHK(launch_rows_c2r<T>(a, st_));
    }
}

// ---------------------------------------------------------------------------
// field I/O
// ---------------------------------------------------------------------------
template <typename T>
void Engine<T>::set_field(int idx, const double* in) {
    if (idx < 0 || idx >= nf_) throw Error(GPF_EINVAL, "field index out of range");
    std::vector<T> h(g_.n);
    double s = 0.0, c = 0.0;
    for (long j = 0; j < g_.n; ++j) {
        const double v = in[2 * j];
        const double t = s + v;
        c += (std::fabs(s) >= std::fabs(v)) ? ((s - t) + v) : ((v - t) + s);
        s = t;
    }
    const double mean = (fp_.ok && split_mean_) ? (s + c) / (double)g_.n : 0.0;
    for (long j = 0; j < g_.n; ++j) h[j] = (T)(in[2 * j] - mean);
    if (fp_.ok) {
        T* tmp = rscratch(0);
        HIPCHK(hipMemcpyAsync(tmp, h.data(), g_.n * sizeof(T), hipMemcpyHostToDevice, st_));
        r2c({tmp}, {S_.get()}, (T)(1.0 / (double)g_.n));
        prime();
        HIPCHK(hipMemcpyAsync(mean_.get(), &mean, sizeof(double), hipMemcpyHostToDevice, st_));
        HIPCHK(hipStreamSynchronize(st_));
    } else {
        HIPCHK(hipMemcpyAsync(F_[idx].get(), h.data(), g_.n * sizeof(T), hipMemcpyHostToDevice, st_));
        HIPCHK(hipStreamSynchronize(st_));
        kvalid_ = false;
    }
}

template <typename T>
void Engine<T>::get_field(int idx, double* out) {
    if (idx < 0 || idx >= nf_) throw Error(GPF_EINVAL, "field index out of range");
    std::vector<T> h(g_.n);
    if (fp_.ok) {
        T* tmp = rscratch(0);
        c2r({S_.get()}, {tmp}, T(1));
        HIPCHK(hipMemcpyAsync(h.data(), tmp, g_.n * sizeof(T), hipMemcpyDeviceToHost, st_));
    } else {
        HIPCHK(hipMemcpyAsync(h.data(), F_[idx].get(), g_.n * sizeof(T), hipMemcpyDeviceToHost, st_));
    }
    HIPCHK(hipStreamSynchronize(st_));
    check_persist();
    const double mean = host_mean();
    for (long j = 0; j < g_.n; ++j) {
        out[2 * j] = (double)h[j] + mean;
        out[2 * j + 1] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// field output (pf/fileIO.go)
// ---------------------------------------------------------------------------
// Float64IO.SaveFields / Uint8IO.SaveFields: <prefix>_<field>_<epoch>.bin for every
// field.  The bytes are packed on the solver's stream; the copy to pinned host memory and
// the file writes overlap the steps that follow.
template <typename T>
void Engine<T>::save_fields(const char* prefix, int epoch, int format) {
    if (format != kIoFloat64 && format != kIoUint8) throw Error(GPF_EINVAL, "unknown field file format");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (!io_) io_.reset(new SnapshotWriter(dev));
    const long n = g_.n;
    const size_t per = (format == kIoFloat64) ? (size_t)n * 8 : (size_t)n;
    if (format == kIoUint8 && !io_mm_.get()) {
        io_part_.alloc(2 * io_blocks(n));
        io_mm_.alloc(2);
    }
    const int slot = io_->acquire(per * nf_);
    char* buf = static_cast<char*>(io_->device_buffer(slot));
    std::vector<std::pair<std::string, std::pair<size_t, size_t>>> files;
    try {
        for (int i = 0; i < nf_; ++i) {
            const T* x;
            const double* mean = nullptr;
            if (fp_.ok) {
                T* tmp = rscratch(0);
                c2r({S_.get()}, {tmp}, T(1));
                x = tmp;
                if (split_mean_) mean = mean_.get();
            } else {
                x = F_[i].get();
            }
            if (format == kIoFloat64) {
                HIPCHK(launch_pack_be64<T>(x, mean, reinterpret_cast<uint64_t*>(buf + i * per), n, st_));
            } else {
                HIPCHK(launch_minmax<T>(x, mean, n, io_part_.get(), io_mm_.get(), st_));
                HIPCHK(launch_pack_u8<T>(x, mean, io_mm_.get(), reinterpret_cast<uint8_t*>(buf + i * per), n, st_));
            }
            files.push_back({std::string(prefix) + "_" + mcopy_.field_names[i] + "_" + std::to_string(epoch) + ".bin",
                             {i * per, per}});
        }
    } catch (...) {
        io_->submit(slot, st_, 0, {});  // releases the slot
        throw;
    }
    io_->submit(slot, st_, per * nf_, std::move(files));
}

template <typename T>
void Engine<T>::flush_io() {
    if (io_) io_->flush();
}

// CsvIO.SaveFields (fileIO.go:103-115): <prefix>_<epoch>.csv through SaveCsv
template <typename T>
void Engine<T>::save_csv(const char* prefix, int epoch) {
    std::vector<std::vector<double>> values(nf_);
    std::vector<double> buf(2 * g_.n);
    for (int i = 0; i < nf_; ++i) {
        get_field(i, buf.data());
        values[i].resize(g_.n);
        for (long j = 0; j < g_.n; ++j) values[i][j] = buf[2 * j];
    }
    std::vector<int> dims(g_.dims, g_.dims + g_.ndim);
    write_csv(std::string(prefix) + "_" + std::to_string(epoch) + ".csv", mcopy_.field_names, values, dims);
}

template <typename T>
double Engine<T>::host_mean() {
    if (!(fp_.ok && split_mean_)) return 0.0;
    double mu = 0.0;
    HIPCHK(hipMemcpyAsync(&mu, mean_.get(), sizeof(double), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
    return mu;
}

template <typename T>
void Engine<T>::energy(int idx, int npoly, const double* poly, double kappa, double* e, double* mass) {
    if (idx < 0 || idx >= nf_) throw Error(GPF_EINVAL, "field index out of range");
    if (npoly > kMaxPoly || npoly < 0) throw Error(GPF_EINVAL, "energy polynomial too long");
    const T* c;
    if (fp_.ok) {
        T* tmp = rscratch(0);
        c2r({S_.get()}, {tmp}, T(1));
        c = tmp;
    } else {
        c = F_[idx].get();
    }
    const long nb = 512;
    Buf<double> pf(nb), px(nb), pg(nb), res(2), dp(kMaxPoly);
    std::vector<double> hp(kMaxPoly, 0.0);
    for (int q = 0; q < npoly; ++q) hp[q] = poly[q];
    HIPCHK(hipMemcpyAsync(dp.get(), hp.data(), kMaxPoly * sizeof(double), hipMemcpyHostToDevice, st_));
    HIPCHK(launch_real_sums<T>(c, g_.n, host_mean(), npoly, dp.get(), pf.get(), px.get(), nb, st_));
    cpx<T>* kc = kscratch(kMaxArr);
    r2c({c}, {kc}, T(1));
    HIPCHK(launch_kgrad_energy<T>(kc, g_.kg, g_.nmain, g_.nk, pg.get(), nb, st_));
    HIPCHK(launch_reduce2(pf.get(), nb, 1.0, pg.get(), nb, 0.5 * kappa / (double)g_.n, res.get(), st_));
    HIPCHK(launch_reduce2(px.get(), nb, 1.0, px.get(), 0, 0.0, res.get() + 1, st_));
    double h[2];
    HIPCHK(hipMemcpyAsync(h, res.get(), 2 * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
    *e = h[0];
    *mass = h[1];
}

// FP64 reductions of field idx as the host reads it: the fused path's real field is the
// row-pass inverse of the state plus its FP64 mean
template <typename T>
void Engine<T>::reduce(int idx, double* out) {
    if (idx < 0 || idx >= nf_) throw Error(GPF_EINVAL, "field index out of range");
    const T* c;
    if (fp_.ok) {
        T* tmp = rscratch(0);
        c2r({S_.get()}, {tmp}, T(1));
        c = tmp;
    } else {
        c = F_[idx].get();
    }
    Buf<double> part(kRedOut * reduce_blocks(g_.n)), res(kRedOut);
    HIPCHK(launch_field_reduce<T>(c, g_.n, (fp_.ok && split_mean_) ? mean_.get() : nullptr, part.get(), res.get(),
                                  st_));
    HIPCHK(hipMemcpyAsync(out, res.get(), kRedOut * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
}

template <typename T>
void Engine<T>::monitor(int npoly, const double* poly, double kappa) {
    if (!fp_.ok) throw Error(GPF_EINVAL, "the per-step energy series is recorded by the fused path only");
    if (npoly > kMaxPoly || npoly < 0) throw Error(GPF_EINVAL, "energy polynomial too long");
    mon_ = true;
    mon_poly_.assign(poly, poly + npoly);
    mon_kappa_ = kappa;
    series_n_ = 0;
    if (!series_count_.get()) series_count_.alloc(1);
    const long zero = 0;
    HIPCHK(hipMemcpyAsync(series_count_.get(), &zero, sizeof(long), hipMemcpyHostToDevice, st_));
    HIPCHK(hipStreamSynchronize(st_));
    drop_graph();
}

template <typename T>
long Engine<T>::series(double* e, double* m, long cap) {
    HIPCHK(hipStreamSynchronize(st_));
    check_persist();
    const long n = std::min(cap, series_n_);
    if (n > 0) {
        std::vector<double> h(2 * series_n_);
        HIPCHK(hipMemcpy(h.data(), series_.get(), 2 * series_n_ * sizeof(double), hipMemcpyDeviceToHost));
        for (long i = 0; i < n; ++i) {
            if (e) e[i] = h[2 * i];
            if (m) m[i] = h[2 * i + 1];
        }
    }
    return series_n_;
}

// ---------------------------------------------------------------------------
// fused single-field step
// ---------------------------------------------------------------------------
template <typename T>
RowFusedArgs<T> Engine<T>::rows_args(cpx<T>* W, double* mean, int tw, int var) {
    RowFusedArgs<T> ra;
    ra.tw = this->tw(g_.nx);
    ra.s_main = W;
    ra.s_nyq = W + g_.nmain;
    ra.g_main = W;
    ra.g_nyq = W + g_.nmain;
    ra.rows = g_.rows;
    ra.nx = g_.nx;
    // untuned: spread over more workgroups on small grids, down to the autotuner's smallest
    // candidate (4 lines per block)
    ra.TW = tw > 0 ? tw : spread_tw(choose_tw(g_.nx, 1, sizeof(cpx<T>), 1, var), (g_.rows + 1) / 2, 4);
    ra.variant = var;
    ra.c_out = c_out_;
    ra.npoly = (int)fp_.gpoly.size();
    for (int q = 0; q < ra.npoly; ++q) ra.gpoly[q] = fp_.gpoly[q];
    ra.mean = split_mean_ ? mean : nullptr;
    return ra;
}

template <typename T>
PencilUpdateArgs<T> Engine<T>::update_args(cpx<T>* S, cpx<T>* W, double* mean, int tw) {
    PencilUpdateArgs<T> u;
    const int axis = upd_axis();
    const int nu = (axis == 2) ? g_.nz : g_.ny;
    u.p.tw = this->tw(nu);
    u.p.narr = 2;
    u.p.arr[0][0] = S;
    u.p.arr[0][1] = S + g_.nmain;
    u.p.arr[1][0] = W;
    u.p.arr[1][1] = W + g_.nmain;
    u.p.seg[0] = (axis == 2) ? seg_z_main(g_) : seg_y_main(g_);
    u.p.seg[1] = (axis == 2) ? seg_z_nyq(g_) : seg_y_nyq(g_);
    u.p.nseg = 2;
    u.p.TW = tw > 0 ? tw
                    : spread_tw(choose_tw(nu, u.p.narr, sizeof(cpx<T>), 8, 1),
                                (long)u.p.seg[0].na * u.p.seg[0].nb + (long)u.p.seg[1].na * u.p.seg[1].nb, 4);
    u.g = g_.kg;
    u.dt = dt_;
    u.scale = 1.0 / (double)g_.n;
    u.denom = fp_.denom;
    u.lin = fp_.lin;
    u.src_lap = fp_.src_lap;
    u.has_src = fp_.has_src ? 1 : 0;
    u.filter = has_filter_ ? filter_.get() : nullptr;
    u.mean = split_mean_ ? mean : nullptr;
    make_dense(u);
    return u;
}

// W <- S inverse-transformed along the update axis: the row-pass input of the next step
template <typename T>
void Engine<T>::prime() {
    pencils({W_.get()}, {S_.get()}, upd_axis(), +1);
}

template <typename T>
void Engine<T>::pencils_tw(const std::vector<cpx<T>*>& dst, int axis, int dir, int tw, int lpt, int var) {
    const int n = (axis == 1) ? g_.ny : g_.nz;
    PencilArgs<T> a;
    a.tw = this->tw(n);
    a.narr = (int)dst.size();
    a.seg[0] = (axis == 1) ? seg_y_main(g_) : seg_z_main(g_);
    a.seg[1] = (axis == 1) ? seg_y_nyq(g_) : seg_z_nyq(g_);
    a.nseg = 2;
    a.TW = tw > 0 ? tw
                  : spread_tw(choose_tw(n, a.narr, sizeof(cpx<T>), 8, var),
                              (long)a.seg[0].na * a.seg[0].nb + (long)a.seg[1].na * a.seg[1].nb, 4);
    a.lpt = lpt;
    a.variant = var;
    for (int q = 0; q < a.narr; ++q) {
        a.arr[q][0] = dst[q];
        a.arr[q][1] = dst[q] + g_.nmain;
    }
    HIPCHK(launch_pencils<T>(a, n, dir, st_));
}

// Lines-per-block autotuner of the fused step: every pass is timed with HIP
// events for each admissible tile width on scratch copies of the state.
template <typename T>
void Engine<T>::autotune() {
    if (!fp_.ok || tuned_) return;
    Buf<cpx<T>> Sb(g_.nk), Wb(g_.nk);
    Buf<double> mu(1);
    HIPCHK(hipMemcpyAsync(Sb.get(), S_.get(), g_.nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
    HIPCHK(hipMemcpyAsync(Wb.get(), W_.get(), g_.nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
    HIPCHK(hipMemcpyAsync(mu.get(), mean_.get(), sizeof(double), hipMemcpyDeviceToDevice, st_));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto time_it = [&](auto&& launch) {
        launch();
        launch();
        HIPCHK(hipEventRecord(e0, st_));
        const int reps = 5;
        for (int r = 0; r < reps; ++r) launch();
        HIPCHK(hipEventRecord(e1, st_));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        return ms / reps;
    };
    const int cand[] = {4, 8, 16, 32, 64};
    auto best_of = [&](int n, auto&& launch_with, int variant = 0) {
        int best = 0;
        float bt = 1e30f;
        for (int tw : cand) {
            if (!tw_fits(n, tw, sizeof(cpx<T>), variant)) continue;
            const float t = time_it([&] { launch_with(tw); });
            if (t < bt) { bt = t; best = tw; }
        }
        return best;
    };
    {
        // row pass: lines per block and transform schedule
        float bt = 1e30f;
        for (int var = 0; var < 2; ++var)
            for (int tw : cand) {
                if (!tw_fits(g_.nx, tw, sizeof(cpx<T>), var)) continue;
                const float t = time_it([&] { HIPCHK(launch_rows_fused<T>(rows_args(Wb.get(), mu.get(), tw, var), st_)); });
                if (t < bt) { bt = t; tw_rows_ = tw; rows_var_ = var; }
            }
    }
    // for each order of the y / z passes, tune the tile widths and keep the faster order
    const int orders = (g_.ndim == 3) ? 2 : 1;
    float best_total = 1e30f;
    int best[4] = {0, 0, 0, 2};
    for (int o = 0; o < orders; ++o) {
        upd_axis_ = (o == 0) ? 2 : 1;
        int pf = 0, pi = 0, pu = 0;
        float total = 0.f;
        const int no = axis_len(other_axis()), nu = axis_len(upd_axis());
        int lpi = 1, lpf = 1, vpi = 0, vpf = 0;
        if (g_.ndim == 3) {
            // schedule, lines per thread and lines per block of the plain pencil passes
            auto best_pencil = [&](int dir, int& tw_out, int& lpt_out, int& var_out) {
                float bt = 1e30f;
                for (int var = 0; var < 2; ++var)
                    for (int lpt = 1; lpt <= (sizeof(T) == 4 ? 2 : 1); ++lpt)
                        for (int tw : cand) {
                            if (!tw_fits(no, tw, sizeof(cpx<T>), var)) continue;
                            const float t = time_it([&] { pencils_tw({Wb.get()}, other_axis(), dir, tw, lpt, var); });
                            if (t < bt) { bt = t; tw_out = tw; lpt_out = lpt; var_out = var; }
                        }
                return bt;
            };
            total += best_pencil(+1, pi, lpi, vpi);
            if (fp_.has_src) total += best_pencil(-1, pf, lpf, vpf);
        }
        pu = best_of(nu, [&](int tw) { HIPCHK(launch_pencil_update<T>(update_args(Sb.get(), Wb.get(), mu.get(), tw), nu, st_)); }, 1);
        total += time_it([&] { HIPCHK(launch_pencil_update<T>(update_args(Sb.get(), Wb.get(), mu.get(), pu), nu, st_)); });
        if (total < best_total) {
            best_total = total;
            best[0] = pf; best[1] = pu; best[2] = pi; best[3] = upd_axis_;
            lpt_pf_ = lpf;
            lpt_pi_ = lpi;
            var_pf_ = vpf;
            var_pi_ = vpi;
        }
    }
    tw_pf_ = best[0];
    tw_pu_ = best[1];
    tw_pi_ = best[2];
    upd_axis_ = best[3];
    HIPCHK(hipEventDestroy(e0));
    HIPCHK(hipEventDestroy(e1));
    HIPCHK(hipStreamSynchronize(st_));
    tuned_ = true;
    warmed_ = true;  // tens of milliseconds of the step's kernels have just run
    prime();  // the work array follows the chosen update axis
}

// TimeStepper.SetFilter (pf/euler.go:62-64, pf/rk4.go:138-140): a modal filter given as its
// 1000-point table, evaluated on the device as ModalFilter.Eval of a pf.Vandeven
// (pf/vandeven.go:31-40); a null table removes the filter
template <typename T>
void Engine<T>::set_filter(const double* table) {
    if (table && stepper_ == GPF_SDD) throw Error(GPF_EINVAL, "SDD: Does not support modal filters");
    HIPCHK(hipStreamSynchronize(st_));
    drop_graph();
    stab_ok_ = false;
    if (!table) {
        has_filter_ = false;
        return;
    }
    if (!filter_.get()) filter_.alloc(1000);
    HIPCHK(hipMemcpy(filter_.get(), table, 1000 * sizeof(double), hipMemcpyHostToDevice));
    has_filter_ = true;
    stab_ok_ = false;
}

template <typename T>
int Engine<T>::tuning(int* tws, int maxn) {
    const int v[10] = {tw_rows_, tw_pf_, tw_pu_, tw_pi_, upd_axis(), rows_var_, lpt_pf_, lpt_pi_, var_pf_, var_pi_};
    for (int i = 0; i < 10 && i < maxn; ++i) tws[i] = v[i];
    return 10;
}

// a launch layout read back with tuning(), applied instead of running the autotuner (the
// same layout on every run of a workload, as FFTW's wisdom)
template <typename T>
void Engine<T>::set_tuning(const int* v, int n) {
    if (!fp_.ok) throw Error(GPF_EINVAL, "launch layouts apply to the fused path only");
    if (n != 10) throw Error(GPF_EINVAL, "a launch layout has 10 values (gpf_solver_tuning order)");
    const int axis = (g_.ndim == 3) ? v[4] : 1;
    if (axis != 1 && axis != 2) throw Error(GPF_EINVAL, "update axis must be 1 (y) or 2 (z)");
    const int other = (axis == 2) ? 1 : 2;
    const int nu = axis == 2 ? g_.nz : g_.ny, no = other == 2 ? g_.nz : g_.ny;
    const size_t el = sizeof(cpx<T>);
    auto ok_tw = [](int tw) { return tw == 4 || tw == 8 || tw == 16 || tw == 32 || tw == 64; };
    const int rows_var = v[5], lpf = v[6], lpi = v[7], vpf = v[8], vpi = v[9];
    bool ok = ok_tw(v[0]) && ok_tw(v[2]) && (rows_var == 0 || rows_var == 1) && tw_fits(g_.nx, v[0], el, rows_var) &&
              tw_fits(nu, v[2], el, 1);
    if (g_.ndim == 3) {
        ok = ok && ok_tw(v[3]) && (vpi == 0 || vpi == 1) && (lpi == 1 || (lpi == 2 && sizeof(T) == 4)) &&
             tw_fits(no, v[3], el, vpi);
        if (fp_.has_src)
            ok = ok && ok_tw(v[1]) && (vpf == 0 || vpf == 1) && (lpf == 1 || (lpf == 2 && sizeof(T) == 4)) &&
                 tw_fits(no, v[1], el, vpf);
    }
    if (!ok) throw Error(GPF_EINVAL, "launch layout not admissible for this grid and precision");
    HIPCHK(hipStreamSynchronize(st_));
    tw_rows_ = v[0];
    tw_pf_ = v[1];
    tw_pu_ = v[2];
    tw_pi_ = v[3];
    upd_axis_ = axis;
    rows_var_ = rows_var;
    lpt_pf_ = lpf;
    lpt_pi_ = lpi;
    var_pf_ = vpf;
    var_pi_ = vpi;
    tuned_ = true;
    drop_graph();
    prime();  // the work array follows the update axis
}

template <typename T>
void Engine<T>::drop_graph() {
    if (graph_) (void)hipGraphExecDestroy(graph_);
    graph_ = nullptr;
    persist_ok_ = false;
}

template <typename T>
void Engine<T>::prepare_fused(long upcoming) {
    if (!mon_) return;
    const long nblk_a = rows_fused_blocks(g_.rows, rows_args(W_.get(), mean_.get(), tw_rows_, rows_var_).TW);
    PencilUpdateArgs<T> u = update_args(S_.get(), W_.get(), mean_.get(), tw_pu_);
    int tiles0 = 0;
    const long nblk_c = pencil_blocks(u.p.seg, u.p.nseg, u.p.TW, &tiles0);
    // two sets of partial sums: consecutive steps alternate (enqueue_fused)
    if (nblk_a_ != nblk_a) { pa_bulk_.alloc(2 * nblk_a); pa_mass_.alloc(2 * nblk_a); nblk_a_ = nblk_a; drop_graph(); }
    if (nblk_c_ != nblk_c) { pc_eg_.alloc(2 * nblk_c); nblk_c_ = nblk_c; drop_graph(); }
    if (series_n_ + upcoming > series_cap_) {
        const long cap = std::max({1024L, 2 * series_cap_, series_n_ + upcoming});
        Buf<double> nb(2 * cap);
        if (series_n_ > 0)
            HIPCHK(hipMemcpyAsync(nb.get(), series_.get(), 2 * series_n_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
        HIPCHK(hipStreamSynchronize(st_));
        series_.swap(nb);
        series_cap_ = cap;
        drop_graph();
    }
}

// With the monitor, step pos writes its partial sums to set pos mod 2, and its energy-series
// slot is written by block 0 of the next step's row launch (the row blocks having at least
// 64 threads); the last step of the sequence writes its own slot with a launch of its own.
template <typename T>
void Engine<T>::enqueue_fused(int pos, int count) {
    RowFusedArgs<T> ra = rows_args(W_.get(), mean_.get(), tw_rows_, rows_var_);
    const int nu = axis_len(upd_axis());
    PencilUpdateArgs<T> u = update_args(S_.get(), W_.get(), mean_.get(), tw_pu_);
    const double ceg = 0.5 * mon_kappa_ / (double)g_.n;
    const long cur = pos & 1, prev = (pos + 1) & 1;
    const bool rows_pass = fp_.has_src || mon_;
    const bool fold = mon_ && rows_pass && pos > 0 && ra.TW * fft_threads_per_line(g_.nx, ra.variant) >= 64;
    if (mon_) {
        ra.nbulk = (int)mon_poly_.size();
        for (int q = 0; q < ra.nbulk; ++q) ra.bulk[q] = mon_poly_[q];
        ra.bulk_partial = pa_bulk_.get() + cur * nblk_a_;
        ra.mass_partial = pa_mass_.get() + cur * nblk_a_;
        u.egrad_partial = pc_eg_.get() + cur * nblk_c_;
    }
    if (fold) {
        ra.prev_bulk = pa_bulk_.get() + prev * nblk_a_;
        ra.prev_mass = pa_mass_.get() + prev * nblk_a_;
        ra.prev_eg = pc_eg_.get() + prev * nblk_c_;
        ra.prev_nb = nblk_a_;
        ra.prev_ne = nblk_c_;
        ra.prev_ceg = ceg;
        ra.series = series_.get();
        ra.counter = series_count_.get();
    } else if (mon_ && pos > 0) {
        // the previous step's slot was not folded into this row launch
        HIPCHK(launch_series_append(pa_bulk_.get() + prev * nblk_a_, nblk_a_, pc_eg_.get() + prev * nblk_c_, nblk_c_,
                                    ceg, pa_mass_.get() + prev * nblk_a_, series_.get(), series_count_.get(), st_));
    }
    // W holds S inverse-transformed along the update axis; the real-space pass is needed
    // for the explicit source and for the monitor's bulk sums
    if (fp_.has_src || mon_) {
        if (g_.ndim == 3) {
            tbegin(3);
            pencils_tw({W_.get()}, other_axis(), +1, tw_pi_, lpt_pi_, var_pi_);
            tend(3);
        }
        tbegin(0);
        HIPCHK(launch_rows_fused<T>(ra, st_));
        tend(0);
        if (g_.ndim == 3 && fp_.has_src) {
            tbegin(1);
            pencils_tw({W_.get()}, other_axis(), -1, tw_pf_, lpt_pf_, var_pf_);
            tend(1);
        }
    }
    tbegin(2);
    HIPCHK(launch_pencil_update<T>(u, nu, st_));
    tend(2);
    if (mon_ && pos == count - 1)
        HIPCHK(launch_series_append(pa_bulk_.get() + cur * nblk_a_, nblk_a_, pc_eg_.get() + cur * nblk_c_, nblk_c_,
                                    ceg, pa_mass_.get() + cur * nblk_a_, series_.get(), series_count_.get(), st_));
}

template <typename T>
void Engine<T>::fused_step() {
    if (autotune_ && !tuned_) {
        autotune();
        drop_graph();
    }
    prepare_fused(1);
    enqueue_fused();
    if (mon_) ++series_n_;
}

template <typename T>
void Engine<T>::capture_graph() {
    if (graph_ && graph_mon_ != mon_) drop_graph();
    if (graph_) return;
    graph_mon_ = mon_;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < kGraphSteps; ++k) enqueue_fused(k, kGraphSteps);
    HIPCHK(hipStreamEndCapture(st_, &g));
    const hipError_t e = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    // the executable graph's device-side copy is made now, not at its first launch
    HIPCHK(hipGraphUpload(graph_, st_));
}

// Runs the step's passes, in the chosen launch layout, on scratch copies of the state
// until ms_min milliseconds have passed (at most 2000 steps): an idle device runs its first
// milliseconds of work at lower clocks, which the autotuner's timing runs otherwise absorb.
template <typename T>
void Engine<T>::warm(double ms_min) {
    if (!fp_.ok) return;
    Buf<cpx<T>> Sb(g_.nk), Wb(g_.nk);
    Buf<double> mu(1);
    HIPCHK(hipMemcpyAsync(Sb.get(), S_.get(), g_.nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
    HIPCHK(hipMemcpyAsync(Wb.get(), W_.get(), g_.nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
    HIPCHK(hipMemcpyAsync(mu.get(), mean_.get(), sizeof(double), hipMemcpyDeviceToDevice, st_));
    const int nu = axis_len(upd_axis());
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < 200; ++k) {
        for (int j = 0; j < 10; ++j) {
            if (fp_.has_src) {
                if (g_.ndim == 3) pencils_tw({Wb.get()}, other_axis(), +1, tw_pi_, lpt_pi_, var_pi_);
                HIPCHK(launch_rows_fused<T>(rows_args(Wb.get(), mu.get(), tw_rows_, rows_var_), st_));
                if (g_.ndim == 3) pencils_tw({Wb.get()}, other_axis(), -1, tw_pf_, lpt_pf_, var_pf_);
            }
            HIPCHK(launch_pencil_update<T>(update_args(Sb.get(), Wb.get(), mu.get(), tw_pu_), nu, st_));
        }
        HIPCHK(hipStreamSynchronize(st_));
        if (std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() >= ms_min) break;
    }
    warmed_ = true;
}

// The autotuner (on scratch copies), the monitor's buffers and the step graph, ahead of
// the first step, so that a caller's timed steps replay graphs only; without the tuner,
// the warm-up on scratch copies brings the device to its steady clocks.  The state is not
// advanced: the tuner's prime() recomputes W from S, which it already holds.
template <typename T>
double Engine<T>::prepare(long upcoming) {
    const auto t0 = std::chrono::steady_clock::now();
    if (fp_.ok) {
        if (autotune_ && !tuned_) {
            autotune();
            drop_graph();
        }
        if (!warmed_) warm(30.0);
        prepare_fused(std::max(upcoming, 0L));
        if (persist_eligible()) {
            if (!persist_ok_) build_persist();
        } else if (!timing_ && !c_out_) {
            capture_graph();
        }
        HIPCHK(hipStreamSynchronize(st_));
    }
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// n fused steps: launch-bound grids replay a captured graph of kGraphSteps steps (kernels
// only: no memset or allocation is captured; the series slot comes from a device counter)
template <typename T>
void Engine<T>::steps(long n) {
    if (!fp_.ok) {
        for (long i = 0; i < n; ++i) step();
        return;
    }
    if (n > 0 && autotune_ && !tuned_) {
        fused_step();
        --n;
    }
    if (n >= 2 && persist_eligible()) {
        prepare_fused(n);
        n = run_persist(n);
        if (n == 0) return;
    }
    if (timing_ || c_out_ || n < kGraphSteps) {
        for (long i = 0; i < n; ++i) fused_step();
        return;
    }
    prepare_fused(n);
    capture_graph();
    while (n >= kGraphSteps) {
        HIPCHK(hipGraphLaunch(graph_, st_));
        n -= kGraphSteps;
        if (mon_) series_n_ += kGraphSteps;
    }
    for (; n > 0; --n) fused_step();
}

// The persistent step runs on square 2D grids of side 32 to 256 (the compiled lengths), with
// the device's cooperative launch, outside per-pass timing and real-space output.  In the
// default mode it is taken where it was measured faster than the replayed per-pass graphs.
template <typename T>
bool Engine<T>::persist_eligible() {
    if (!fp_.ok || persist_mode_ == 0 || timing_ || c_out_ || g_.ndim != 2 || g_.nx != g_.ny || !persist_side(g_.nx))
        return false;
    // by default only with the participants on one XCD (the agent-fenced barrier is slower
    // than the per-pass launches)
    if (persist_mode_ < 0 && (g_.nx > kPersistAutoMax || local_ == 0)) return false;
    if (coop_ < 0) {
        int dev = 0, v = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeCooperativeLaunch, dev));
        coop_ = v ? 1 : 0;
    }
    return coop_ == 1;
}

template <typename T>
void Engine<T>::set_persistent(int mode, int threads) {
    if (mode < -1 || mode > 2)
        throw Error(GPF_EINVAL, "persistent mode must be -1 (auto), 0 (off), 1 (on) or 2 (on, agent-fenced barrier)");
    if (threads != 0 && threads != 64 && threads != 128 && threads != 256 && threads != 512)
        throw Error(GPF_EINVAL, "threads per block of the persistent step: 0 (64), 64, 128, 256 or 512");
    if (threads > 0 && g_.ndim == 2 && persist_side(g_.nx) && g_.nx == g_.ny) {
        const int t1 = fft_threads_per_line(g_.nx, 1);
        if (threads < t1 || !tw_fits(g_.nx, threads / t1, sizeof(cpx<T>), 1))
            throw Error(GPF_EINVAL, "threads per block not admissible for this grid's persistent step");
    }
    HIPCHK(hipStreamSynchronize(st_));
    persist_mode_ = mode;
    persist_threads_ = threads;
    persist_ok_ = false;
}

// the plan: the row pass in schedule 1 and the update pass with the same threads per block,
// their partial sums in two sets (consecutive steps alternate)
template <typename T>
void Engine<T>::build_persist() {
    const int n = g_.nx;
    const int t1 = fft_threads_per_line(n, 1);
    const int thr = persist_threads_ > 0 ? persist_threads_ : std::max(64, t1);
    const int tw = thr / t1;
    if (tw < 1 || tw * t1 != thr || !tw_fits(n, tw, sizeof(cpx<T>), 1))
        throw Error(GPF_EINVAL, "persistent step: no admissible block for this grid");
    PersistPlan<T> P;
    RowFusedArgs<T> ra = rows_args(W_.get(), mean_.get(), tw, 1);
    PencilUpdateArgs<T> u = update_args(S_.get(), W_.get(), mean_.get(), tw);
    u.p.n = n;
    u.p.lpt = 1;
    int tiles0 = 0;
    const long nr = rows_fused_blocks(g_.rows, tw);
    const long nu = pencil_blocks(u.p.seg, u.p.nseg, tw, &tiles0);
    u.p.tiles0 = tiles0;
    // the multipliers are the same every step: tabulated once, read with the lines
    if (!ptab_.get()) ptab_.alloc(g_.nk);
    if (!pm0_.get()) pm0_.alloc(2);
    u.mtab[0] = ptab_.get();
    u.mtab[1] = ptab_.get() + g_.nmain;
    u.m0 = pm0_.get();
    HIPCHK(launch_update_mult_table<T>(u, st_));
    P.rows_pass = (fp_.has_src || mon_) ? 1 : 0;
    P.nblk_rows = (int)nr;
    P.nblk_upd = (int)nu;
    for (int k = 0; k < 3; ++k) P.ra[k] = ra;
    P.u[0] = P.u[1] = u;
    if (mon_) {
        if ((long)ppa_bulk_.size() != 2 * nr) { ppa_bulk_.alloc(2 * nr); ppa_mass_.alloc(2 * nr); }
        if ((long)ppc_eg_.size() != 2 * nu) ppc_eg_.alloc(2 * nu);
        const double ceg = 0.5 * mon_kappa_ / (double)g_.n;
        for (int k = 0; k < 3; ++k) {
            RowFusedArgs<T>& a = P.ra[k];
            const long cur = (k == 1) ? 1 : 0, prev = 1 - cur;
            a.nbulk = (int)mon_poly_.size();
            for (int q = 0; q < a.nbulk; ++q) a.bulk[q] = mon_poly_[q];
            a.bulk_partial = ppa_bulk_.get() + cur * nr;
            a.mass_partial = ppa_mass_.get() + cur * nr;
            if (k > 0) {
                a.prev_bulk = ppa_bulk_.get() + prev * nr;
                a.prev_mass = ppa_mass_.get() + prev * nr;
                a.prev_eg = ppc_eg_.get() + prev * nu;
                a.prev_nb = nr;
                a.prev_ne = nu;
                a.prev_ceg = ceg;
                a.series = series_.get();
                a.counter = series_count_.get();
            }
        }
        P.u[0].egrad_partial = ppc_eg_.get();
        P.u[1].egrad_partial = ppc_eg_.get() + nu;
        if (P.rows_pass && nu > nr) {
            // the participant beyond the row blocks writes the slots
            for (int k = 1; k < 3; ++k) P.ra[k].series = nullptr;
            P.slot_series = series_.get();
            P.slot_counter = series_count_.get();
        }
    }
    if (!pbar_.get()) pbar_.alloc(4);
    if (!pfail_) {
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&pfail_), sizeof(int), hipHostMallocMapped));
        *pfail_ = 0;
        HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&pfail_dev_), pfail_, 0));
    }
    pl_ = PersistLaunch<T>{};
    pl_.plan = P;
    pl_.n = n;
    pl_.mon = mon_ ? 1 : 0;
    persist_geometry<T>(pl_, tw);
    pl_.blocks = (int)(P.rows_pass ? std::max(nr, nu) : nu);
    pl_.bar = pbar_.get();
    pl_.spin_limit = 1u << 20;
    pl_.local = (local_ != 0 && persist_mode_ != 2) ? 1 : 0;
    pl_.host_fail = pfail_dev_;
    pnr_ = nr;
    pnu_ = nu;
    persist_ok_ = true;
}

// n steps in launches of at most 2^20 steps; with the monitor, the slot of a launch's last
// step is written by a launch of its own after it.  The first launch with the participants on
// one XCD is waited for: when the device placed them otherwise (no step run), the launch
// layout falls back to the agent-fenced barrier (mode 1) or to the per-pass launches.
template <typename T>
long Engine<T>::run_persist(long n) {
    check_persist();
    if (!persist_ok_) build_persist();
    const double ceg = 0.5 * mon_kappa_ / (double)g_.n;
    while (n > 0) {
        const int k = (int)std::min(n, 1L << 20);
        HIPCHK(hipMemsetAsync(pbar_.get(), 0, 4 * sizeof(unsigned), st_));
        pl_.steps = k;
        HIPCHK(launch_fused_persist<T>(pl_, st_));
        if (pl_.local && local_ < 0) {
            HIPCHK(hipStreamSynchronize(st_));
            if (*reinterpret_cast<volatile int*>(pfail_) == 2) {
                *pfail_ = 0;
                local_ = 0;
                pl_.local = 0;
                if (persist_mode_ < 0) return n;
                continue;
            }
            local_ = 1;
        }
        if (mon_) {
            const long cur = (k - 1) & 1;
            HIPCHK(launch_series_append(ppa_bulk_.get() + cur * pnr_, pnr_, ppc_eg_.get() + cur * pnu_, pnu_, ceg,
                                        ppa_mass_.get() + cur * pnr_, series_.get(), series_count_.get(), st_));
            series_n_ += k;
        }
        persist_steps_ += k;
        n -= k;
    }
    return 0;
}

template <typename T>
void Engine<T>::check_persist() {
    const int f = pfail_ ? *reinterpret_cast<volatile int*>(pfail_) : 0;
    if (f == 0) return;
    *pfail_ = 0;
    if (f == 2)
        throw Error(GPF_EDEVICE, "persistent step: the device placed the workgroups on more than one XCD; the steps "
                                 "of that launch were not run");
    throw Error(GPF_EDEVICE, "persistent step: a device-wide barrier wait exceeded its limit; the fields of that run "
                             "are not valid");
}

// ---------------------------------------------------------------------------
// general path
// ---------------------------------------------------------------------------
template <typename T>
void Engine<T>::ensure_k_fields() {
    if (kvalid_) return;
    std::vector<const T*> in;
    std::vector<cpx<T>*> out;
    for (int i = 0; i < nf_; ++i) {
        in.push_back(F_[i].get());
        out.push_back(K_[i].get());
    }
    r2c(in, out, T(1));
    kvalid_ = true;
}

template <typename T>
void Engine<T>::sync_derived() {
    if (nd_ == 0) return;
    std::vector<const T*> in;
    std::vector<cpx<T>*> out;
    // dense programs (polynomials of at most two fields) run inside the row transform;
    // others (non-integer powers, noise) are evaluated into Dr_ first
    std::vector<const DerivedArgs<T>*> prog;
    for (int d = 0; d < nd_; ++d) {
        dargs_[d].counter = dcount_[d]++;
        const bool inline_prog = dargs_[d].dense && !dargs_[d].noise;
        if (!inline_prog) HIPCHK(launch_derived<T>(dargs_[d], st_));
        in.push_back(inline_prog ? nullptr : Dr_[d].get());
        prog.push_back(inline_prog ? &dargs_[d] : nullptr);
        out.push_back(K_[nf_ + d].get());
    }
    r2c(in, out, T(1), &prog);
}

// SquaredGradient terms of equation eq, evaluated on the current k-space bricks
template <typename T>
void Engine<T>::eval_user_terms_for(int eq) {
    for (const auto& t : eqs_[eq].terms) {
        if (t.kind != kSquaredGradient) continue;
        const UserTerm& ut = mcopy_.user_table[kSquaredGradient][t.user];
        const int b = mcopy_.brick_index(ut.field);
        const int nd = g_.ndim;
        KGradArgs<T> a;
        a.src = K_[b].get();
        for (int d = 0; d < nd; ++d) a.out[d] = kscratch(kMaxArr + 1 + d);
        a.g = g_.kg;
        a.nmain = g_.nmain;
        a.nk = g_.nk;
        HIPCHK(launch_kgrad<T>(a, nd, st_));
        std::vector<const cpx<T>*> in;
        std::vector<T*> out;
        for (int d = 0; d < nd; ++d) {
            in.push_back(a.out[d]);
            out.push_back(rscratch(1 + d));
        }
        c2r(in, out, T(1) / (T)g_.n);
        const T* w[3] = {out[0], out[nd > 1 ? 1 : 0], out[nd > 2 ? 2 : 0]};
        T* sq = rscratch(4);
        HIPCHK(launch_sqsum<T>(w, nd, (T)ut.factor, sq, g_.n, st_));
        r2c({sq}, {Tsq_[t.user].get()}, T(1));
    }
}

// the scalar kernel K(k) of elastic term e: material C and sigma* = C : misfit
template <typename T>
KElasticArgs<T> Engine<T>::elastic_args(size_t e) const {
    const UserTerm& ut = mcopy_.user_table[kHomogeneousModulus][e];
    KElasticArgs<T> a;
    std::memcpy(a.ep.C, ut.C, sizeof(a.ep.C));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) s = s + ut.C[i * 27 + j * 9 + k * 3 + l] * ut.misfit[k * 3 + l];
            a.ep.sig[i * 3 + j] = s;
        }
    a.ep.dim = g_.ndim;
    a.g = g_.kg;
    a.nmain = g_.nmain;
    a.nk = g_.nk;
    return a;
}

// HomogeneousModulusLinElast: FFT(H'(phi) (IFFT(K H^) / N - 2 e* H(phi))), once per step.
// The indicator H(phi) = 3 phi^2 - 2 phi^3 is evaluated inside the row transform of phi
// (a dense one-field polynomial), and the inverse row transform of K H^, the pointwise
// H'(phi) (S - 2 e* H(phi)) and its forward row transform are one k_rows_sync pass, so
// neither H, S nor the real source goes through HBM.
template <typename T>
void Engine<T>::eval_elastic() {
    for (size_t e = 0; e < Tel_.size(); ++e) {
        const UserTerm& ut = mcopy_.user_table[kHomogeneousModulus][e];
        cpx<T>* Hk = kscratch(kMaxArr + 1);
        DerivedArgs<T> ind;
        ind.nf = 1;
        ind.f[0] = lag_[e].get();
        ind.dense = 1;
        ind.dc[2][0] = 3.0;
        ind.dc[3][0] = -2.0;
        const std::vector<const DerivedArgs<T>*> prog{&ind};
        r2c({nullptr}, {Hk}, T(1), &prog);
        KElasticArgs<T> a = elastic_args(e);
        a.src = Hk;
        a.dst = Hk;
        a.ktab = Ktab_[e].get();
        HIPCHK(launch_kelastic<T>(a, st_));
        if (g_.ndim == 3) pencils({Hk}, {}, 2, +1);
        pencils({Hk}, {}, 1, +1);
        RowSyncArgs<T> r;
        r.tw = tw(g_.nx);
        r.fmain[0] = Hk;
        r.fnyq[0] = Hk + g_.nmain;
        r.scale = T(1) / (T)g_.n;
        r.mode = 1;
        r.phi = lag_[e].get();
        r.two_estar = 2.0 * energy_density(ut.C, ut.misfit);
        r.nd = 1;
        r.dmain[0] = Tel_[e].get();
        r.dnyq[0] = Tel_[e].get() + g_.nmain;
        r.rows = g_.rows;
        r.nx = g_.nx;
        r.TW = spread_tw(choose_tw(g_.nx, 1, sizeof(cpx<T>), 1), g_.rows);
        HIPCHK(launch_rows_sync<T>(r, st_));
        pencils({Tel_[e].get()}, {}, 1, -1);
        if (g_.ndim == 3) pencils({Tel_[e].get()}, {}, 2, -1);
    }
}

// dst = filt * (alpha * cur + beta * (sum_t terms_t [+ extra])) / (1 - gamma * D)
template <typename T>
void Engine<T>::rhs_update(int eq, cpx<T>* dst, const cpx<T>* cur, double alpha, double beta, double gamma,
                           bool filt, const cpx<T>* extra) {
    KUpdateArgs<T> a;
    a.dst = dst;
    a.cur = cur;
    a.g = g_.kg;
    a.nmain = g_.nmain;
    a.nk = g_.nk;
    a.alpha = alpha;
    a.beta = beta;
    a.gamma = gamma;
    a.denom = eqs_[eq].denom;
    a.sv = eqs_[eq].sv;
    a.hs = eqs_[eq].hs;
    a.filter = (filt && has_filter_) ? filter_.get() : nullptr;
    if (extra) {
        a.brick[0] = extra;
        a.coef[0] = 1.0;
        a.pw[0] = 0;
        a.nterm = 1;
    } else {
        terms_into(eq, a);
    }
    HIPCHK(launch_kupdate<T>(a, st_));
}

// the explicit terms of equation eq (GetRHS, model.go:283-303) as k-space bricks
template <typename T>
template <class A>
void Engine<T>::terms_into(int eq, A& a) {
    a.nterm = 0;
    for (const auto& t : eqs_[eq].terms) {
        if (t.kind == kConservativeNoise) {
            // sum_comp i 2 sin(pi f_comp) FFT(current_comp) (noise.go:60-78)
            for (size_t c = 0; c < cn_bricks_[t.user].size(); ++c) {
                if (a.nterm >= kMaxUTerms) throw Error(GPF_EINVAL, "too many terms in one equation");
                const int q = a.nterm++;
                a.coef[q] = t.coef;
                a.pw[q] = t.pw;
                a.dmul[q] = nullptr;
                a.tmode[q] = 1;
                a.tcomp[q] = (int)c;
                a.brick[q] = K_[cn_bricks_[t.user][c]].get();
            }
            continue;
        }
        if (a.nterm >= kMaxUTerms) throw Error(GPF_EINVAL, "too many terms in one equation");
        const int q = a.nterm++;
        a.coef[q] = t.coef;
        a.pw[q] = t.pw;
        a.dmul[q] = nullptr;
        a.tmode[q] = 0;
        if (t.kind == kConcrete) {
            a.brick[q] = (t.brick >= 0) ? K_[t.brick].get() : nullptr;
        } else if (t.kind == kSquaredGradient) {
            a.brick[q] = Tsq_[t.user].get();
        } else if (t.kind == kVolumeConserving) {
            // Multiplier * indicator brick (volumeConserving.go:19-28)
            a.brick[q] = K_[vc_brick_[t.user]].get();
            a.dmul[q] = vc_state_[t.user].get();
        } else {
            a.brick[q] = Tel_[t.user].get();
        }
    }
}

template <typename T>
void Engine<T>::general_step_euler() {
    ensure_k_fields();
    sync_derived();
    eval_elastic();
    for (int i = 0; i < nf_; ++i) {
        eval_user_terms_for(i);
        rhs_update(i, K_[i].get(), K_[i].get(), 1.0, dt_, dt_, true, nullptr);
    }
    std::vector<const cpx<T>*> in;
    std::vector<T*> out;
    for (int i = 0; i < nf_; ++i) {
        in.push_back(K_[i].get());
        out.push_back(F_[i].get());
    }
    c2r(in, out, T(1) / (T)g_.n);
}

// Stages 0-2 of RK4: the corrected fields' inverse transform and the derived fields'
// forward transform meet in one row pass (k_rows_sync), so the real fields are not written
// (the last stage writes them).  Used when every derived field is a dense polynomial of
// the (at most two) fields.
template <typename T>
void Engine<T>::sync_from_spectra(const PencilStageArgs<T>* stage) {
    std::vector<cpx<T>*> scr;
    std::vector<const cpx<T>*> src;
    for (int i = 0; i < nf_; ++i) {
        scr.push_back(kscratch(i));
        src.push_back(K_[i].get());
    }
    if (g_.ndim == 3) {
        pencils(scr, src, 2, +1, stage);
        pencils(scr, {}, 1, +1);
    } else {
        pencils(scr, src, 1, +1, stage);
    }
    RowSyncArgs<T> a;
    a.tw = tw(g_.nx);
    for (int i = 0; i < nf_; ++i) {
        a.fmain[i] = scr[i];
        a.fnyq[i] = scr[i] + g_.nmain;
    }
    a.scale = T(1) / (T)g_.n;
    a.nd = nd_;
    std::vector<cpx<T>*> out;
    for (int d = 0; d < nd_; ++d) {
        dargs_[d].counter = dcount_[d]++;
        std::memcpy(a.dc[d], dargs_[d].dc, sizeof(a.dc[d]));
        a.dmain[d] = K_[nf_ + d].get();
        a.dnyq[d] = K_[nf_ + d].get() + g_.nmain;
        out.push_back(K_[nf_ + d].get());
    }
    a.rows = g_.rows;
    a.nx = g_.nx;
    a.TW = spread_tw(choose_tw(g_.nx, 1, sizeof(cpx<T>), 1), g_.rows);
    HIPCHK(launch_rows_sync<T>(a, st_));
    pencils(out, {}, 1, -1);
    if (g_.ndim == 3) pencils(out, {}, 2, -1);
}

// The factor 1 / (1 - gamma dt D(k)) of RK4's stages depends only on the plan and on gamma
// (1/2 for stages 0-1, 1 for stages 2-3, the filter on the last): tabulated once in FP64 and
// stored in T, so the stage kernels read it instead of forming it in every stage.
template <typename T>
void Engine<T>::build_stage_tables() {
    stab_.clear();
    for (int i = 0; i < nf_; ++i) {
        bool ok = true;
        for (const auto& t : eqs_[i].terms) ok &= (t.kind != kConservativeNoise);
        for (int k = 0; k < 3; ++k) {
            stab_.emplace_back();
            if (!ok || (k == 2 && !has_filter_)) continue;
            stab_.back().alloc(g_.nk);
            StageDenomArgs<T> a;
            a.denom = eqs_[i].denom;
            a.sv = eqs_[i].sv;
            a.hs = eqs_[i].hs;
            a.gamma = (k == 0) ? 0.5 * dt_ : dt_;
            a.filter = (k == 2) ? filter_.get() : nullptr;
            a.g = g_.kg;
            a.nmain = g_.nmain;
            a.nk = g_.nk;
            a.out = stab_.back().get();
            HIPCHK(launch_stage_denoms<T>(a, st_));
        }
    }
    stab_ok_ = true;
}

// A stage's update done by the fields' first inverse pencil pass (launch_pencils_stage)
// instead of its own pass: when every field's equation has at most kStageTerms terms of
// tmode 0, none formed per stage (SquaredGradient), a stage factor table, and the first
// inverse axis has a core length.  The pass reads the operands the stage kernel reads and
// does its arithmetic, so the result is the same; the corrected spectrum is stored only
// after the last stage (the next step's state) or when an equation reads a field's
// spectrum as a brick.
template <typename T>
bool Engine<T>::fused_stage_args(int stage, PencilStageArgs<T>* out) {
    if (sizeof(T) != 8 || !stab_ok_ || nf_ > kMaxArr) return false;  // FP64 only (spectral_stage.hip)
    if (!stage_pencil_length(g_.ndim == 3 ? g_.nz : g_.ny)) return false;
    const double fac_prep[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const double fac_corr[3] = {0.5, 0.5, 1.0};
    PencilStageArgs<T> a;
    a.stage = stage;
    a.dt = dt_;
    a.fp = fac_prep[stage];
    a.fc = stage < 3 ? fac_corr[stage] : 0.0;
    a.g = g_.kg;
    a.nmain = g_.nmain;
    a.nk = g_.nk;
    bool field_brick = false;
    for (int i = 0; i < nf_; ++i) {
        for (const auto& t : eqs_[i].terms)
            if (t.kind == kSquaredGradient || t.kind == kConservativeNoise) return false;
        Rk4StageArgs<T> r;
        terms_into(i, r);
        if (r.nterm > kStageTerms) return false;
        const int kind = (stage < 2) ? 0 : (stage == 3 && has_filter_) ? 2 : 1;
        StageLine<T>& f = a.f[i];
        f.stab = stab_[3 * i + kind].get();
        if (!f.stab) return false;
        f.k = K_[i].get();
        f.init = INIT_[i].get();
        f.final = FINAL_[i].get();
        f.kn = KF_[i].get();
        f.nterm = r.nterm;
        for (int q = 0; q < r.nterm; ++q) {
            if (r.tmode[q] != 0) return false;
            f.brick[q] = r.brick[q];
            f.coef[q] = r.coef[q];
            f.pw[q] = r.pw[q];
            f.dmul[q] = r.dmul[q];
            a.need_l |= (r.pw[q] != 0) ? 1 : 0;
            for (int j = 0; j < nf_; ++j) field_brick |= (r.brick[q] == K_[j].get());
        }
    }
    a.store_kn = (stage == 3 || field_brick) ? 1 : 0;
    *out = a;
    return true;
}

template <typename T>
void Engine<T>::general_step_rk4() {
    // RK4.Step (rk4.go:29-127) with each stage's GetRHS, PrepareNextCorrection and
    // correction fused into one k-space pass per field: the stage's right-hand side is
    // evaluated from the bricks inside the pass instead of being stored.  The corrected
    // spectra go to KF_ (free, as the right-hand side is never stored) and swap with K_
    // after all fields, so every field of a stage reads the same bricks.  Where possible
    // (fused_stage_args) that pass is the fields' first inverse pencil pass itself.
    if (!stab_ok_) build_stage_tables();
    ensure_k_fields();
    sync_derived();
    eval_elastic();  // the lagged field is constant within a step (homoLinElast.go:130-134)
    const double fac_prep[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const double fac_corr[3] = {0.5, 0.5, 1.0};
    for (int stage = 0; stage < 4; ++stage) {
        PencilStageArgs<T> fs;
        const bool fused = fused_stage_args(stage, &fs);
        for (int i = 0; i < nf_ && !fused; ++i) {
            eval_user_terms_for(i);
            Rk4StageArgs<T> a;
            terms_into(i, a);
            a.k = K_[i].get();
            a.kn = KF_[i].get();
            a.init = INIT_[i].get();
            a.final = FINAL_[i].get();
            a.stage = stage;
            a.dt = dt_;
            a.fp = fac_prep[stage];
            a.fc = stage < 3 ? fac_corr[stage] : 0.0;
            a.denom = eqs_[i].denom;
            a.sv = eqs_[i].sv;
            a.hs = eqs_[i].hs;
            a.filter = (stage == 3 && has_filter_) ? filter_.get() : nullptr;
            a.g = g_.kg;
            a.nmain = g_.nmain;
            a.nk = g_.nk;
            const int kind = (stage < 2) ? 0 : (stage == 3 && has_filter_) ? 2 : 1;
            a.stab = stab_[3 * i + kind].get();
            a.need_l = 0;
            for (int t = 0; t < a.nterm; ++t) a.need_l |= (a.pw[t] != 0);
            HIPCHK(launch_rk4_stage<T>(a, st_));
        }
        std::vector<const cpx<T>*> in;
        std::vector<T*> out;
        for (int i = 0; i < nf_; ++i) {
            K_[i].swap(KF_[i]);
            in.push_back(K_[i].get());
            out.push_back(F_[i].get());
        }
        if (stage < 3 && row_sync_) {
            sync_from_spectra(fused ? &fs : nullptr);
            continue;
        }
        c2r(in, out, T(1) / (T)g_.n, nullptr, fused ? &fs : nullptr);
        if (stage < 3) sync_derived();
    }
}

template <typename T>
double Engine<T>::dev_scalar(const double* p) {
    double v = 0.0;
    HIPCHK(hipMemcpyAsync(&v, p, sizeof(double), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
    return v;
}

template <typename T>
void Engine<T>::implicit_residual(const double* x, double* out) {
    if constexpr (std::is_same<T, double>::value) {
        const long n = g_.n;
        std::vector<const T*> in;
        std::vector<cpx<T>*> ko;
        if (row_sync_ && nf_ + nd_ <= kMaxArr) {
            // the fields' transform and the dense derived fields' (their polynomials of x
            // evaluated in the row pass) in one row launch and one launch per pencil axis;
            // the real fields are not needed until the step takes the solution
            std::vector<DerivedArgs<T>> px(nd_);
            std::vector<const DerivedArgs<T>*> prog;
            for (int i = 0; i < nf_; ++i) {
                in.push_back(x + i * n);
                ko.push_back(K_[i].get());
                prog.push_back(nullptr);
            }
            for (int d = 0; d < nd_; ++d) {
                dargs_[d].counter = dcount_[d]++;
                px[d] = dargs_[d];
                for (int f = 0; f < nf_; ++f) px[d].f[f] = x + f * n;
                in.push_back(nullptr);
                ko.push_back(K_[nf_ + d].get());
                prog.push_back(&px[d]);
            }
            r2c(in, ko, T(1), &prog);
        } else {
            for (int i = 0; i < nf_; ++i) {
                HIPCHK(launch_copy<T>(x + i * n, F_[i].get(), n, st_));
                in.push_back(F_[i].get());
                ko.push_back(K_[i].get());
            }
            r2c(in, ko, T(1));
            sync_derived();
        }
        std::vector<const cpx<T>*> rin;
        std::vector<T*> rout;
        for (int i = 0; i < nf_; ++i) {
            eval_user_terms_for(i);
            EtdArgs<T> a;
            terms_into(i, a);
            a.dst = RES_[i].get();
            a.cur = K_[i].get();
            a.orig = ORIG_[i].get();
            a.rprev = RP_[i].get();
            a.denom = eqs_[i].denom;
            a.sv = eqs_[i].sv;
            a.hs = eqs_[i].hs;
            a.dt = dt_;
            a.g = g_.kg;
            a.nmain = g_.nmain;
            a.nk = g_.nk;
            HIPCHK(launch_etd_residual<T>(a, st_));
            rin.push_back(RES_[i].get());
            rout.push_back(out + i * n);
        }
        c2r(rin, rout, T(1) / (T)n);
    } else {
        (void)x;
        (void)out;
        throw Error(GPF_EINVAL, "ImplicitEuler runs in FP64 (GPF_FP64)");
    }
}

// central difference weights of the first derivative on 2, 4, 6, 8 points (Fornberg)
struct FdStencil {
    int n;
    int off[8];
    double w[8];
};
static const FdStencil kFdStencils[4] = {
    {2, {-1, 1}, {-0.5, 0.5}},
    {4, {-2, -1, 1, 2}, {1.0 / 12.0, -2.0 / 3.0, 2.0 / 3.0, -1.0 / 12.0}},
    {6, {-3, -2, -1, 1, 2, 3}, {-1.0 / 60.0, 3.0 / 20.0, -3.0 / 4.0, 3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0}},
    {8, {-4, -3, -2, -1, 1, 2, 3, 4},
     {1.0 / 280.0, -4.0 / 105.0, 1.0 / 5.0, -4.0 / 5.0, 4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0}},
};

// out = J(x) v for the current iterate x (in nx_, with F(x) in nfx_)
template <typename T>
void Engine<T>::implicit_jv(const double* v, double* out) {
    const long M = (long)nf_ * g_.n;
    const double h = nk_.step;
    if (nk_.stencil == 1) {
        HIPCHK(launch_axpby(1.0, nx_.get(), h, v, nxp_.get(), M, st_));
        implicit_residual(nxp_.get(), nfp_.get());
        HIPCHK(launch_axpby(1.0 / h, nfp_.get(), -1.0 / h, nfx_.get(), out, M, st_));
        return;
    }
    const FdStencil& fd = kFdStencils[nk_.stencil / 2 - 1];
    for (int k = 0; k < fd.n; ++k) {
        HIPCHK(launch_axpby(1.0, nx_.get(), fd.off[k] * h, v, nxp_.get(), M, st_));
        implicit_residual(nxp_.get(), nfp_.get());
        HIPCHK(launch_axpby(1.0, k == 0 ? nullptr : out, fd.w[k] / h, nfp_.get(), out, M, st_));
    }
}

// restarted GMRES for J dx = -F(x), dx = 0 initially; returns the Arnoldi steps taken
template <typename T>
int Engine<T>::implicit_gmres() {
    const long M = (long)nf_ * g_.n;
    const int m = nk_.restart;
    const long nb = krylov_blocks(M);
    constexpr int kStride = kMaxKrylov + 1;
    double* V = nv_.get();
    double* dev = ndev_.get();  // [0, kStride) CGS pass 1, [kStride, 2 kStride) pass 2, [2 kStride] |w|^2,
                                // [3 kStride, 4 kStride) solution coefficients
    HIPCHK(hipMemsetAsync(ndx_.get(), 0, M * sizeof(double), st_));
    HIPCHK(launch_axpby(0.0, nullptr, -1.0, nfx_.get(), nw_.get(), M, st_));
    HIPCHK(launch_multidot(nw_.get(), M, 1, nw_.get(), M, npart_.get(), st_));
    HIPCHK(launch_reduce_rows(npart_.get(), 1, nb, dev + 2 * kStride, st_));
    HIPCHK(launch_normalize(nw_.get(), dev + 2 * kStride, V, M, st_));
    double beta = std::sqrt(std::max(dev_scalar(dev + 2 * kStride), 0.0));
    if (!(beta > 0.0) || !std::isfinite(beta)) return 0;
    // the linear model needs no residual far below the Newton tolerance (max norm <= 2-norm)
    const double target = std::max(1e-10 * beta, 0.1 * nk_.tol);
    std::vector<double> H((size_t)(m + 1) * m), Hu((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m + 1);
    auto at = [m](std::vector<double>& A, int i, int j) -> double& { return A[(size_t)i * m + j]; };
    nhost_.assign(2 * kStride + 1, 0.0);
    int total = 0;
    const int max_cycles = 10;
    for (int cycle = 0; cycle < max_cycles; ++cycle) {
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(Hu.begin(), Hu.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int jend = 0;
        bool done = false;
        for (int j = 0; j < m; ++j) {
            double* w = nw_.get();
            implicit_jv(V + (long)j * M, w);
            // classical Gram-Schmidt, twice
            HIPCHK(launch_multidot(V, M, j + 1, w, M, npart_.get(), st_));
            HIPCHK(launch_reduce_rows(npart_.get(), j + 1, nb, dev, st_));
            HIPCHK(launch_combine(V, M, j + 1, dev, -1.0, w, M, st_));
            HIPCHK(launch_multidot(V, M, j + 1, w, M, npart_.get(), st_));
            HIPCHK(launch_reduce_rows(npart_.get(), j + 1, nb, dev + kStride, st_));
            HIPCHK(launch_combine(V, M, j + 1, dev + kStride, -1.0, w, M, st_));
            HIPCHK(launch_multidot(w, M, 1, w, M, npart_.get(), st_));
            HIPCHK(launch_reduce_rows(npart_.get(), 1, nb, dev + 2 * kStride, st_));
            HIPCHK(launch_normalize(w, dev + 2 * kStride, V + (long)(j + 1) * M, M, st_));
            HIPCHK(hipMemcpyAsync(nhost_.data(), dev, (2 * kStride + 1) * sizeof(double), hipMemcpyDeviceToHost,
                                  st_));
            HIPCHK(hipStreamSynchronize(st_));
            for (int i = 0; i <= j; ++i) at(Hu, i, j) = at(H, i, j) = nhost_[i] + nhost_[kStride + i];
            const double hn = std::sqrt(std::max(nhost_[2 * kStride], 0.0));
            at(Hu, j + 1, j) = at(H, j + 1, j) = hn;
            // Givens rotations of the least-squares problem min |g - H y|
            for (int i = 0; i < j; ++i) {
                const double a = at(H, i, j), b = at(H, i + 1, j);
                at(H, i, j) = cs[i] * a + sn[i] * b;
                at(H, i + 1, j) = -sn[i] * a + cs[i] * b;
            }
            const double a = at(H, j, j), b = at(H, j + 1, j);
            const double r = std::hypot(a, b);
            cs[j] = (r > 0.0) ? a / r : 1.0;
            sn[j] = (r > 0.0) ? b / r : 0.0;
            at(H, j, j) = r;
            at(H, j + 1, j) = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++total;
            jend = j + 1;
            if (!std::isfinite(g[j + 1])) return total;
            if (std::fabs(g[j + 1]) <= target || !(hn > 1e-14 * beta)) {
                done = true;
                break;
            }
        }
        // back substitution for y, then dx += V y
        for (int i = jend - 1; i >= 0; --i) {
            double t = g[i];
            for (int k = i + 1; k < jend; ++k) t -= at(H, i, k) * y[k];
            y[i] = (at(H, i, i) != 0.0) ? t / at(H, i, i) : 0.0;
        }
        HIPCHK(hipMemcpyAsync(dev + 3 * kStride, y.data(), jend * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPCHK(launch_combine(V, M, jend, dev + 3 * kStride, 1.0, ndx_.get(), M, st_));
        HIPCHK(hipStreamSynchronize(st_));
        if (done || cycle + 1 == max_cycles) break;
        // restart from the Arnoldi residual r = V_{m+1} (beta e1 - Hu y)
        std::vector<double> z(jend + 1, 0.0);
        z[0] = beta;
        for (int i = 0; i <= jend; ++i)
            for (int k = 0; k < jend; ++k) z[i] -= at(Hu, i, k) * y[k];
        HIPCHK(hipMemcpyAsync(dev + 3 * kStride, z.data(), (jend + 1) * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPCHK(hipMemsetAsync(nw_.get(), 0, M * sizeof(double), st_));
        HIPCHK(launch_combine(V, M, jend + 1, dev + 3 * kStride, 1.0, nw_.get(), M, st_));
        HIPCHK(launch_multidot(nw_.get(), M, 1, nw_.get(), M, npart_.get(), st_));
        HIPCHK(launch_reduce_rows(npart_.get(), 1, nb, dev + 2 * kStride, st_));
        HIPCHK(launch_normalize(nw_.get(), dev + 2 * kStride, V, M, st_));
        beta = std::sqrt(std::max(dev_scalar(dev + 2 * kStride), 0.0));
        if (!(beta > target) || !std::isfinite(beta)) break;
    }
    return total;
}

template <typename T>
void Engine<T>::implicit_step() {
    const long n = g_.n;
    const long M = (long)nf_ * n;
    if (!nv_.get() || nv_.size() != (size_t)(nk_.restart + 1) * M) nv_.alloc((size_t)(nk_.restart + 1) * M);
    // orig^ and R^prev at the current fields (implicitEuler.go:172-179)
    ensure_k_fields();
    sync_derived();
    eval_elastic();  // the lagged elastic field is constant within a step
    const size_t kb = g_.nk * sizeof(cpx<T>);
    for (int i = 0; i < nf_; ++i) {
        eval_user_terms_for(i);
        rhs_update(i, RP_[i].get(), nullptr, 0.0, 1.0, 0.0, false, nullptr);
        HIPCHK(hipMemcpyAsync(ORIG_[i].get(), K_[i].get(), kb, hipMemcpyDeviceToDevice, st_));
    }
    // explicit Euler step as the initial guess (implicitEuler.go:182-190; euler.go:32-44 with
    // the stepper's filter), from the right-hand side just formed
    std::vector<const cpx<T>*> in;
    std::vector<T*> out;
    for (int i = 0; i < nf_; ++i) {
        rhs_update(i, K_[i].get(), K_[i].get(), 1.0, dt_, dt_, true, RP_[i].get());
        in.push_back(K_[i].get());
        out.push_back(F_[i].get());
    }
    c2r(in, out, T(1) / (T)n);
    for (int i = 0; i < nf_; ++i)
        HIPCHK(hipMemcpyAsync(nx_.get() + i * n, F_[i].get(), n * sizeof(T), hipMemcpyDeviceToDevice, st_));
    // Newton iteration (nonlin.NewtonKrylov.Solve)
    nk_newton_ = 0;
    nk_krylov_ = 0;
    nk_conv_ = 0;
    double* dev = ndev_.get();
    for (int it = 0; it < nk_.maxiter; ++it) {
        implicit_residual(nx_.get(), nfx_.get());
        HIPCHK(launch_absmax(nfx_.get(), M, npart_.get(), dev, st_));
        nk_res_ = dev_scalar(dev);
        if (nk_res_ < nk_.tol) {
            nk_conv_ = 1;
            break;
        }
        if (!std::isfinite(nk_res_)) break;
        nk_krylov_ += implicit_gmres();
        HIPCHK(launch_axpby(1.0, nx_.get(), 1.0, ndx_.get(), nx_.get(), M, st_));
        ++nk_newton_;
    }
    if (!nk_conv_) std::fprintf(stderr, "Warning: Iterative solver did not converge\n");
    // the fields take the solution (implicitEuler.go:205); after a converged check the
    // k-space bricks of the fields are FFT(x) already
    for (int i = 0; i < nf_; ++i)
        HIPCHK(hipMemcpyAsync(F_[i].get(), nx_.get() + i * n, n * sizeof(T), hipMemcpyDeviceToDevice, st_));
    kvalid_ = nk_conv_ != 0;
}

template <typename T>
void Engine<T>::set_newton_krylov(int maxiter, double step_size, double tol, int stencil, int restart) {
    if (maxiter < 1) throw Error(GPF_EINVAL, "Maxiter must be positive");
    if (!(step_size > 0.0) || !(tol > 0.0)) throw Error(GPF_EINVAL, "StepSize and Tol must be positive");
    if (stencil != 1 && stencil != 2 && stencil != 4 && stencil != 6 && stencil != 8)
        throw Error(GPF_EINVAL, "Stencil must be 1, 2, 4, 6 or 8");
    if (restart < 1 || restart > kMaxKrylov) throw Error(GPF_EINVAL, "GMRES restart must be in 1..64");
    nk_.maxiter = maxiter;
    nk_.step = step_size;
    nk_.tol = tol;
    nk_.stencil = stencil;
    nk_.restart = restart;
}

template <typename T>
void Engine<T>::newton_stats(int* newton_iters, int* krylov_iters, int* converged, double* residual) const {
    if (newton_iters) *newton_iters = nk_newton_;
    if (krylov_iters) *krylov_iters = nk_krylov_;
    if (converged) *converged = nk_conv_;
    if (residual) *residual = nk_res_;
}

// ---------------------------------------------------------------------------
// SDD: shrinking-dimer dynamics (sdd.go:84-460)
// ---------------------------------------------------------------------------
// Step (sdd.go:159-301) on the device: the right-hand sides at the dimer's ends x -+ l v / 2,
// the Householder-reflected force on the centre with the Sherman-Morrison solve of the
// implicit part, the torque on the orientation and its explicit update.  Every full-spectrum
// sum of the reference becomes a Hermitian-weighted half-spectrum sum; the scalars stay in
// device memory, so a step needs no host synchronisation.  As in the reference, every
// field's weighted force is built from the first field's right-hand sides (sdd.go:212-216).
template <typename T>
double Engine<T>::sdd_dimer_length(double t) const {
    const double l = sdd_len0_ * std::exp(-t / sdd_tau_l_);
    return (l < sdd_min_len_) ? sdd_min_len_ : l;
}

template <typename T>
void Engine<T>::sdd_init(const double* init, const double* final, long n) {
    if (stepper_ != GPF_SDD) throw Error(GPF_EINVAL, "the solver's stepper is not SDD");
    if (n != (long)nf_ * g_.n) throw Error(GPF_EINVAL, "buffer length does not match nfields * nodes");
    std::vector<double> v(n);
    for (long i = 0; i < n; ++i) v[i] = final[i] - init[i];
    sdd_set_orientation(v.data(), n);
}

template <typename T>
void Engine<T>::sdd_set_orientation(const double* v, long n) {
    if (stepper_ != GPF_SDD) throw Error(GPF_EINVAL, "the solver's stepper is not SDD");
    if (n != (long)nf_ * g_.n) throw Error(GPF_EINVAL, "Inconsistent length of the passed orientaiton vector");
    double len = 0.0;
    for (long i = 0; i < n; ++i) len += v[i] * v[i];
    len = std::sqrt(len);
    std::vector<double> u(n);
    for (long i = 0; i < n; ++i) u[i] = v[i] / len;
    HIPCHK(hipMemcpy(sdd_v_.get(), u.data(), n * sizeof(double), hipMemcpyHostToDevice));
    sdd_len0_ = len;
    sdd_ready_ = true;
}

template <typename T>
void Engine<T>::sdd_set_params(double alpha, double tau_o, double tau_l, double init_len, double min_len) {
    if (stepper_ != GPF_SDD) throw Error(GPF_EINVAL, "the solver's stepper is not SDD");
    sdd_alpha_ = alpha;
    sdd_tau_o_ = tau_o;
    sdd_tau_l_ = tau_l;
    if (init_len >= 0.0) sdd_len0_ = init_len;
    sdd_min_len_ = min_len;
}

template <typename T>
void Engine<T>::sdd_monitor(double* out) {
    if (stepper_ != GPF_SDD) throw Error(GPF_EINVAL, "the solver's stepper is not SDD");
    HIPCHK(hipMemcpyAsync(out, sdd_mon_.get(), 6 * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
}

template <typename T>
void Engine<T>::sdd_orientation(double* out, long n) {
    if (stepper_ != GPF_SDD) throw Error(GPF_EINVAL, "the solver's stepper is not SDD");
    if (n != (long)nf_ * g_.n) throw Error(GPF_EINVAL, "buffer length does not match nfields * nodes");
    HIPCHK(hipMemcpyAsync(out, sdd_v_.get(), n * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
}

template <typename T>
void Engine<T>::sdd_step() {
    if constexpr (std::is_same<T, double>::value) {
        if (!sdd_ready_) throw Error(GPF_ESTATE, "SDD: The method have to be initialized first. See SDD.Init");
        if (dt_ < 1e-16) throw Error(GPF_EINVAL, "Timestep not set in SDD.");
        const long n = g_.n;
        const long M = (long)nf_ * n;
        const double inv_n = 1.0 / (double)n;
        const double len = sdd_dimer_length((double)sdd_step_ * dt_);
        double* part = sdd_part_.get();
        double* sc = sdd_sc_.get();
        double* mon = sdd_mon_.get();
        // FieldNorm (sdd.go:371-380) before the step
        const long nb = krylov_blocks(n);
        for (int i = 0; i < nf_; ++i) {
            HIPCHK(launch_multidot(F_[i].get(), n, 1, F_[i].get(), n, npart_sdd(), st_));
            HIPCHK(launch_reduce_rows(npart_sdd(), 1, nb, sc + 8 + i, st_));
        }
        HIPCHK(launch_sdd_scalars(6, sc + 8, sc, mon, (double)nf_, 0.0, st_));
        eval_elastic();  // the lagged elastic field is constant within a step
        // orientation spectra
        {
            std::vector<const T*> in;
            std::vector<cpx<T>*> out;
            for (int i = 0; i < nf_; ++i) {
                in.push_back(sdd_v_.get() + i * n);
                out.push_back(sdd_ov_[i].get());
            }
            r2c(in, out, T(1));
        }
        auto shift = [&](double scale) {
            for (int i = 0; i < nf_; ++i)
                HIPCHK(launch_axpby(1.0, F_[i].get(), scale, sdd_v_.get() + i * n, F_[i].get(), n, st_));
        };
        auto transform = [&]() {  // SDD.fft: SyncDerivedFields, then FFT of the fields
            sync_derived();
            std::vector<const T*> in;
            std::vector<cpx<T>*> out;
            for (int i = 0; i < nf_; ++i) {
                in.push_back(F_[i].get());
                out.push_back(K_[i].get());
            }
            r2c(in, out, T(1));
        };
        auto rhs = [&](std::vector<Buf<cpx<T>>>& dst) {
            for (int i = 0; i < nf_; ++i) {
                eval_user_terms_for(i);
                rhs_update(i, dst[i].get(), nullptr, 0.0, 1.0, 0.0, false, nullptr);
            }
        };
        shift(-0.5 * len);
        transform();
        rhs(sdd_rs_);
        shift(len);
        transform();
        rhs(sdd_re_);
        shift(-0.5 * len);
        transform();
        auto args = [&](int i) {
            SddArgs a;
            a.g = g_.kg;
            a.nmain = g_.nmain;
            a.nk = g_.nk;
            a.inv_n = inv_n;
            a.denom = eqs_[i].denom;
            a.sv = eqs_[i].sv;
            a.hs = eqs_[i].hs;
            a.dt = dt_;
            a.alpha = sdd_alpha_;
            a.len = len;
            a.rs0 = sdd_rs_[0].get();
            a.re0 = sdd_re_[0].get();
            a.ov = sdd_ov_[i].get();
            a.k = K_[i].get();
            a.torque = sdd_rs_[i].get();
            a.re = sdd_re_[i].get();
            a.sc = sc;
            a.part = part;
            return a;
        };
        // the centre's update, field by field (sdd.go:205-245)
        for (int i = 0; i < nf_; ++i) {
            const SddArgs a = args(i);
            HIPCHK(launch_sdd_dot(a, st_));
            HIPCHK(launch_sdd_scalars(0, part, sc, mon, inv_n, 0.0, st_));
            HIPCHK(launch_sdd_sm(a, st_));
            HIPCHK(launch_sdd_scalars(1, part, sc, mon, 0.0, 0.0, st_));
            HIPCHK(launch_sdd_update(a, st_));
            HIPCHK(launch_sdd_scalars(i == 0 ? 5 : 2, part, sc, mon, 0.0, 0.0, st_));
        }
        // torque (sdd.go:252-266): rs - re - D ov l, projected off the orientation, to real space
        for (int i = 0; i < nf_; ++i) HIPCHK(launch_sdd_torque(args(i), i == 0, st_));
        HIPCHK(launch_sdd_scalars(3, part, sc, mon, inv_n, 0.0, st_));
        std::vector<const cpx<T>*> tin;
        std::vector<T*> tout;
        for (int i = 0; i < nf_; ++i) {
            HIPCHK(launch_sdd_project(args(i), st_));
            tin.push_back(sdd_rs_[i].get());
            tout.push_back(sdd_tau_.get() + i * n);
        }
        c2r(tin, tout, T(1) / (T)n);
        // orientation (sdd.go:277-291)
        HIPCHK(launch_sdd_orient(sdd_v_.get(), sdd_tau_.get(), M, dt_ / (len * sdd_tau_o_), part, st_));
        HIPCHK(launch_sdd_scalars(4, part, sc, mon, 0.0, (double)M, st_));
        HIPCHK(launch_normalize(sdd_v_.get(), sc + 4, sdd_v_.get(), M, st_));
        // the fields back to real space (sdd.go:293)
        std::vector<const cpx<T>*> in;
        std::vector<T*> out;
        for (int i = 0; i < nf_; ++i) {
            in.push_back(K_[i].get());
            out.push_back(F_[i].get());
        }
        c2r(in, out, T(1) / (T)n);
        kvalid_ = true;
        ++sdd_step_;
    } else {
        throw Error(GPF_EINVAL, "SDD runs in FP64 (GPF_FP64)");
    }
}

template <typename T>
void Engine<T>::step() {
    if (fp_.ok) {
        fused_step();
        return;
    }
    tbegin(0);
    if (stepper_ == GPF_EULER) general_step_euler();
    else if (stepper_ == GPF_RK4) general_step_rk4();
    else if (stepper_ == GPF_SDD) sdd_step();
    else implicit_step();
    tend(0);
}

// OnStepFinished of every registered term (solver.go:73-83): only
// HomogeneousModulusLinElast keeps state, the real part of its field (homoLinElast.go:130-134)
template <typename T>
void Engine<T>::on_step_finished() {
    if (fp_.ok) return;
    // VolumeConservingLP (volumeConserving.go:31-50): the field integral over its first
    // NumNodes nodes (real space after the step); the indicator integral is the zero mode of
    // the derived brick (still in k-space), or element 0 of a field brick
    for (size_t v = 0; v < vc_state_.size(); ++v) {
        const UserTerm& ut = mcopy_.user_table[kVolumeConserving][v];
        const long nb = 512;
        HIPCHK(launch_real_sums<T>(F_[vc_field_[v]].get(), ut.num_nodes, 0.0, 0, nullptr, vc_part_.get(),
                                   vc_part_.get() + nb, nb, st_));
        const int b = vc_brick_[v];
        const T* ind = (b < nf_) ? F_[b].get() : reinterpret_cast<const T*>(K_[b].get());
        HIPCHK(launch_volcons_update<T>(vc_part_.get() + nb, nb, ind, ut.dt, vc_state_[v].get(), st_));
    }
    for (size_t e = 0; e < lag_.size(); ++e) {
        const UserTerm& ut = mcopy_.user_table[kHomogeneousModulus][e];
        int f = 0;
        for (int i = 0; i < nf_; ++i)
            if (mcopy_.field_names[i] == ut.field) f = i;
        HIPCHK(launch_copy<T>(F_[f].get(), lag_[e].get(), g_.n, st_));
    }
}

// ---------------------------------------------------------------------------
// adjoint of the fused Euler step (no reference counterpart)
// ---------------------------------------------------------------------------
// The fused class advances  c^_{n+1} = m1 c^_n + m2 F[g(c_n)]  with real, even
// multipliers m1 = A (1 + dt E), m2 = A dt L^p, A = sigma / (1 - dt D) (euler.go:32-44
// with the fused plan's D, E, g).  Each multiplier is a symmetric operator on real
// fields, so the transposed step is  lam_n = F^-1[m1 F lam_{n+1}] + g'(c_n) F^-1[m2 F lam_{n+1}].
// The parameter gradient of step n is <lam_{n+1}, d c_{n+1} / d s>: the parts through D and E
// are summed over the half spectrum with Hermitian weights against the stored state
// spectra, the part through the source polynomial as the real-space product
// <F^-1[m2 F lam_{n+1}], dg/ds(c_n)> (Parseval).  States are recomputed per segment from
// checkpoints of the state spectrum taken every `every` steps; the recompute keeps each
// spectrum and the real field its row pass writes, so a reverse step costs one forward
// and one inverse transform.
template <typename T>
void Engine<T>::adjoint(long nsteps, long every, const double* target, double* loss, double* grad_field, int nscal,
                        const char* const* names, double* grad_scalars) {
    if (!fp_.ok) throw Error(GPF_EINVAL, "the adjoint covers the fused single-field Euler class (" + fp_.why + ")");
    if (nsteps < 1 || every < 1 || !target || !loss) throw Error(GPF_EINVAL, "bad adjoint arguments");
    if (nscal < 0 || nscal > kMaxAdjScalars) throw Error(GPF_EINVAL, "at most 8 scalars per adjoint run");
    if (nscal > 0 && (!names || !grad_scalars)) throw Error(GPF_EINVAL, "scalar names and output are required");
    if (autotune_ && !tuned_) autotune();
    const bool mon = mon_, timing = timing_;
    mon_ = false;
    timing_ = false;  // the segments are captured as graphs: no per-pass events
    const long n = g_.n, nk = g_.nk;
    // parameter tables: derivative of D, E (per LAP power) and of g (per power of c)
    std::vector<double> dcoef((size_t)std::max(nscal, 1) * 2 * kMaxKTerms, 0.0);
    std::vector<std::vector<double>> dpoly(nscal);
    bool need_c1 = false, need_c0 = false;
    for (int s = 0; s < nscal; ++s) {
        const std::string name = names[s] ? names[s] : "";
        auto it = mcopy_.scalars.find(name);
        if (it == mcopy_.scalars.end()) throw Error(GPF_EINVAL, "unknown scalar " + name);
        if (it->second.imag() != 0.0) throw Error(GPF_EINVAL, "complex scalar " + name);
        FusedPlan d = compile_fused(mcopy_, stepper_, [&](const TermDesc& t, double* c) {
            *c = coef_derivative(t, mcopy_.scalars, name);
            return true;
        });
        if (!d.ok || d.denom.n != fp_.denom.n || d.lin.n != fp_.lin.n)
            throw Error(GPF_ESTATE, "derivative plan does not match the fused plan");
        for (int t = 0; t < d.denom.n; ++t) {
            dcoef[(size_t)s * 2 * kMaxKTerms + t] = d.denom.coef[t];
            need_c1 |= d.denom.coef[t] != 0.0;
        }
        for (int t = 0; t < d.lin.n; ++t) {
            dcoef[(size_t)s * 2 * kMaxKTerms + kMaxKTerms + t] = d.lin.coef[t];
            need_c0 |= d.lin.coef[t] != 0.0;
        }
        if (fp_.has_src) dpoly[s] = d.gpoly;
    }
    // g'(c) coefficients
    std::vector<double> dg;
    if (fp_.has_src)
        for (size_t q = 1; q < fp_.gpoly.size(); ++q) dg.push_back((double)q * fp_.gpoly[q]);
    Buf<double> d_dcoef(dcoef.size()), d_dg(kMaxPoly), d_dpoly((size_t)kMaxPoly * std::max(nscal, 1));
    {
        std::vector<double> h(kMaxPoly, 0.0), hp((size_t)kMaxPoly * std::max(nscal, 1), 0.0);
        for (size_t q = 0; q < dg.size(); ++q) h[q] = dg[q];
        for (int s = 0; s < nscal; ++s)
            for (size_t q = 0; q < dpoly[s].size() && q < (size_t)kMaxPoly; ++q) hp[(size_t)s * kMaxPoly + q] = dpoly[s][q];
        HIPCHK(hipMemcpyAsync(d_dcoef.get(), dcoef.data(), dcoef.size() * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPCHK(hipMemcpyAsync(d_dg.get(), h.data(), kMaxPoly * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPCHK(hipMemcpyAsync(d_dpoly.get(), hp.data(), hp.size() * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPCHK(hipStreamSynchronize(st_));
    }
    // forward sweep with checkpoints of the fused state
    const long nck = (nsteps + every - 1) / every;
    std::vector<Buf<cpx<T>>> ck;
    Buf<double> ckmean(nck);
    for (long s = 0; s < nsteps; s += every) {
        ck.emplace_back(nk);
        HIPCHK(hipMemcpyAsync(ck.back().get(), S_.get(), nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
        HIPCHK(hipMemcpyAsync(ckmean.get() + s / every, mean_.get(), sizeof(double), hipMemcpyDeviceToDevice, st_));
        steps(std::min(every, nsteps - s));
    }
    Buf<cpx<T>> Sf(nk);
    Buf<double> meanf(1);
    HIPCHK(hipMemcpyAsync(Sf.get(), S_.get(), nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
    HIPCHK(hipMemcpyAsync(meanf.get(), mean_.get(), sizeof(double), hipMemcpyDeviceToDevice, st_));
    // loss and lam_T = 2 (c_T - target), in FP64 on the host
    std::vector<double> cT(2 * n);
    get_field(0, cT.data());
    std::vector<T> lamT(n);
    double js = 0.0, jc = 0.0;
    for (long i = 0; i < n; ++i) {
        const double r = cT[2 * i] - target[i];
        const double v = r * r;
        const double t = js + v;
        jc += (std::fabs(js) >= std::fabs(v)) ? ((js - t) + v) : ((v - t) + js);
        js = t;
        lamT[i] = (T)(2.0 * r);
    }
    *loss = js + jc;
    Buf<T> lam(n), pr(n), qr(n);
    HIPCHK(hipMemcpyAsync(lam.get(), lamT.data(), n * sizeof(T), hipMemcpyHostToDevice, st_));
    // segment storage: the state spectra S_n (and means) and the real fields c_n that the
    // recomputed row passes write as a by-product
    const long seglen = std::min(every, nsteps);
    const bool need_real = fp_.has_src;
    std::vector<Buf<cpx<T>>> SS;
    std::vector<Buf<T>> R;
    for (long i = 0; i <= seglen; ++i) SS.emplace_back(nk);
    if (need_real)
        for (long i = 0; i < seglen; ++i) R.emplace_back(n);
    Buf<double> segmean(seglen + 1);
    bool any_g = false;
    for (int q = 0; q < nscal; ++q) any_g |= !dpoly[q].empty();
    Buf<cpx<T>> Lk(nk), Pk(nk), Qk(nk);
    const long nblk = std::max(1L, std::min((nk + 255) / 256, 1024L));
    const long nblk_r = std::max(1L, std::min((n + 255) / 256, 1024L));
    // per-block partial sums, accumulated in place across the reverse sweep (each block owns
    // its slot; launches are ordered on the stream) and reduced once at the end
    Buf<double> partial_k((size_t)std::max(nscal, 1) * nblk), partial_r((size_t)std::max(nscal, 1) * nblk_r),
        acc(std::max(nscal, 1));
    HIPCHK(hipMemsetAsync(acc.get(), 0, std::max(nscal, 1) * sizeof(double), st_));
    HIPCHK(hipMemsetAsync(partial_k.get(), 0, std::max(nscal, 1) * nblk * sizeof(double), st_));
    HIPCHK(hipMemsetAsync(partial_r.get(), 0, std::max(nscal, 1) * nblk_r * sizeof(double), st_));
    AdjArgs<T> a;
    a.g = g_.kg;
    a.nmain = g_.nmain;
    a.nk = nk;
    a.dt = dt_;
    a.denom = fp_.denom;
    a.lin = fp_.lin;
    a.src_lap = fp_.src_lap;
    a.filter = has_filter_ ? filter_.get() : nullptr;
    a.lam = Lk.get();
    a.P = Pk.get();
    a.Q = Qk.get();
    a.nscal = nscal;
    a.dcoef = d_dcoef.get();
    a.cscale = (double)n;
    a.partial = partial_k.get();
    a.inv_n = 1.0 / (double)n;
    const int ndg = (int)dg.size();
    // one segment's recompute and reverse sweep; device pointers depend only on the step's
    // offset in its segment, so a full segment can be captured once and replayed (copies
    // are kernels, nothing is allocated)
    auto copy_k = [&](cpx<T>* dst, const cpx<T>* src) {
        HIPCHK(launch_copy<T>(reinterpret_cast<const T*>(src), reinterpret_cast<T*>(dst), 2 * nk, st_));
    };
    auto recompute = [&](long len) {
        for (long k = 0; k < len; ++k) {
            copy_k(SS[k].get(), S_.get());
            if (split_mean_) HIPCHK(launch_copy<double>(mean_.get(), segmean.get() + k, 1, st_));
            c_out_ = need_real ? R[k].get() : nullptr;
            fused_step();
            c_out_ = nullptr;
        }
        copy_k(SS[len].get(), S_.get());
        if (split_mean_) HIPCHK(launch_copy<double>(mean_.get(), segmean.get() + len, 1, st_));
    };
    // Unless a scalar's gradient needs the source-derivative sums <q, dg/ds(c_n)> (kept in
    // k_adj_lambda), the inverse rows of P and Q, lam_n = p + g'(c_n) q and the forward
    // rows of lam_n are one k_rows_sync pass (mode 2), whose output is the next reverse
    // step's FFT(lam); the first comes from lam_T before the sweep.
    const bool fuse_rev = !(nscal > 0 && any_g) && ndg <= kMaxPoly;
    auto reverse = [&](long len) {
        for (long k = len - 1; k >= 0; --k) {
            const T* cn = need_real ? R[k].get() : nullptr;
            if (!fuse_rev) r2c({lam.get()}, {Lk.get()}, T(1));
            if (nscal > 0 && (need_c1 || need_c0)) {
                // P, Q and the derivatives of D and E against the transforms of c_{n+1} and c_n
                a.C1 = need_c1 ? SS[k + 1].get() : nullptr;
                a.C0 = need_c0 ? SS[k].get() : nullptr;
                a.mean1 = split_mean_ ? segmean.get() + (k + 1) : nullptr;
                a.mean0 = split_mean_ ? segmean.get() + k : nullptr;
                HIPCHK(launch_adj_grad<T>(a, nblk, st_));
            } else {
                HIPCHK(launch_adj_mult<T>(a, st_));
            }
            if (fuse_rev) {
                if (g_.ndim == 3) pencils({Pk.get(), Qk.get()}, {}, 2, +1);
                pencils({Pk.get(), Qk.get()}, {}, 1, +1);
                RowSyncArgs<T> r;
                r.tw = tw(g_.nx);
                r.fmain[0] = Pk.get();
                r.fnyq[0] = Pk.get() + g_.nmain;
                r.fmain[1] = Qk.get();
                r.fnyq[1] = Qk.get() + g_.nmain;
                r.fout[0] = lam.get();
                r.scale = T(1) / (T)n;
                r.mode = 2;
                r.phi = cn;
                r.npoly = ndg;
                for (int q = 0; q < ndg; ++q) r.gd[q] = dg[q];
                r.nd = 1;
                r.dmain[0] = Lk.get();
                r.dnyq[0] = Lk.get() + g_.nmain;
                r.rows = g_.rows;
                r.nx = g_.nx;
                r.TW = spread_tw(choose_tw(g_.nx, 1, sizeof(cpx<T>), 1), g_.rows);
                HIPCHK(launch_rows_sync<T>(r, st_));
                pencils({Lk.get()}, {}, 1, -1);
                if (g_.ndim == 3) pencils({Lk.get()}, {}, 2, -1);
                continue;
            }
            c2r({Pk.get(), Qk.get()}, {pr.get(), qr.get()}, T(1) / (T)n);
            if (need_real) {
                // lam_n = p + g'(c_n) q, and the source derivatives <q, dg/ds(c_n)> in real space
                HIPCHK(launch_adj_lambda<T>(pr.get(), qr.get(), cn, ndg, d_dg.get(), lam.get(), n, any_g ? nscal : 0,
                                            d_dpoly.get(), partial_r.get(), nblk_r, st_));
            } else {
                HIPCHK(launch_copy<T>(pr.get(), lam.get(), n, st_));
            }
        }
    };
    auto capture = [&](auto&& body) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        HIPCHK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
        body();
        HIPCHK(hipStreamEndCapture(st_, &g));
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
        return ge;
    };
    if (fuse_rev) r2c({lam.get()}, {Lk.get()}, T(1));
    hipGraphExec_t g_rec = nullptr, g_rev = nullptr;
    bool warm = false;  // the first segment runs eagerly: it creates every lazily built table
    for (long seg = nck - 1; seg >= 0; --seg) {
        const long s0 = seg * every, s1 = std::min(nsteps, s0 + every), len = s1 - s0;
        HIPCHK(hipMemcpyAsync(S_.get(), ck[seg].get(), nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
        HIPCHK(hipMemcpyAsync(mean_.get(), ckmean.get() + seg, sizeof(double), hipMemcpyDeviceToDevice, st_));
        prime();
        if (warm && len == every && len >= 2) {
            if (!g_rec) g_rec = capture([&] { recompute(len); });
            HIPCHK(hipGraphLaunch(g_rec, st_));
            if (!g_rev) g_rev = capture([&] { reverse(len); });
            HIPCHK(hipGraphLaunch(g_rev, st_));
        } else {
            recompute(len);
            reverse(len);
        }
        warm = true;
    }
    if (nscal > 0) {
        HIPCHK(launch_accumulate(partial_k.get(), nscal, nblk, acc.get(), st_));
        HIPCHK(launch_accumulate(partial_r.get(), nscal, nblk_r, acc.get(), st_));
    }
    HIPCHK(hipStreamSynchronize(st_));
    if (g_rec) (void)hipGraphExecDestroy(g_rec);
    if (g_rev) (void)hipGraphExecDestroy(g_rev);
    // outputs; the solver is left at c_T
    std::vector<T> hl(n);
    HIPCHK(hipMemcpyAsync(hl.data(), lam.get(), n * sizeof(T), hipMemcpyDeviceToHost, st_));
    std::vector<double> hs(std::max(nscal, 1));
    HIPCHK(hipMemcpyAsync(hs.data(), acc.get(), hs.size() * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipMemcpyAsync(S_.get(), Sf.get(), nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
    HIPCHK(hipMemcpyAsync(mean_.get(), meanf.get(), sizeof(double), hipMemcpyDeviceToDevice, st_));
    prime();
    HIPCHK(hipStreamSynchronize(st_));
    if (grad_field)
        for (long i = 0; i < n; ++i) grad_field[i] = (double)hl[i];
    for (int s = 0; s < nscal; ++s) grad_scalars[s] = hs[s];
    mon_ = mon;
    timing_ = timing;
}

// ---------------------------------------------------------------------------
// full complex FourierTransform (pfutil.FFTWWrapper)
// ---------------------------------------------------------------------------
class FFTBase {
public:
    virtual ~FFTBase() = default;
    virtual void exec(double* data, int dir) = 0;
    virtual long size() const = 0;
};

// the longest length the FourierTransform drop-in transforms without a compiled schedule
// (Bluestein with a convolution of at most 2048 points)
constexpr int kBluesteinMax = 1024;

template <typename T>
class FullFFT : public FFTBase {
public:
    FullFFT(const int* dims, int ndim) : ndim_(ndim) {
        if (ndim < 1 || ndim > 3) throw Error(GPF_EINVAL, "FourierTransform needs 1 to 3 dimensions");
        nx_ = dims[ndim - 1];
        ny_ = (ndim >= 2) ? dims[ndim - 2] : 1;
        nz_ = (ndim == 3) ? dims[0] : 1;
        for (int v : {nx_, ny_, nz_})
            if (v > 1 && !fft_size_supported(v) && v > kBluesteinMax)
                throw Error(GPF_EINVAL, "unsupported transform length " + std::to_string(v) +
                                            " (lengths without a compiled schedule must be <= 1024)");
        n_ = (long)nx_ * ny_ * nz_;
    }
    ~FullFFT() override {
        if (st_) (void)hipStreamDestroy(st_);
    }
    long size() const override { return n_; }
    void exec(double* data, int dir) override {
        ensure();
        std::vector<cpx<T>> h(n_);
        for (long i = 0; i < n_; ++i) {
            h[i].x = (T)data[2 * i];
            h[i].y = (T)data[2 * i + 1];
        }
        HIPCHK(hipMemcpyAsync(d_.get(), h.data(), n_ * sizeof(cpx<T>), hipMemcpyHostToDevice, st_));
        if (nx_ > 1) {
            Seg s;
            s.na = 1; s.nb = (int)(ny_ * nz_); s.sb = nx_; s.es = 1;
            pass(s, nx_, dir);
        }
        if (ny_ > 1) {
            Seg s;
            s.na = nz_; s.nb = nx_; s.sa = (long)ny_ * nx_; s.sb = 1; s.es = nx_;
            pass(s, ny_, dir);
        }
        if (nz_ > 1) {
            Seg s;
            s.na = ny_; s.nb = nx_; s.sa = nx_; s.sb = 1; s.es = (long)nx_ * ny_;
            pass(s, nz_, dir);
        }
        HIPCHK(hipMemcpyAsync(h.data(), d_.get(), n_ * sizeof(cpx<T>), hipMemcpyDeviceToHost, st_));
        HIPCHK(hipStreamSynchronize(st_));
        for (long i = 0; i < n_; ++i) {
            data[2 * i] = (double)h[i].x;
            data[2 * i + 1] = (double)h[i].y;
        }
    }

private:
    // device resources on first use: the plan itself (NewFFTW) is host data, so a
    // length mismatch is reported before anything touches the device
    void ensure() {
        if (st_) return;
        HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        d_.alloc(n_);
        auto upload_tw = [&](int v) {
            for (auto& p : tw_)
                if (p.first == v) return;
            auto h = twiddle_host<T>(v);
            Buf<cpx<T>> b(v);
            HIPCHK(hipMemcpy(b.get(), h.data(), v * sizeof(cpx<T>), hipMemcpyHostToDevice));
            tw_.emplace_back(v, std::move(b));
        };
        for (int v : {nx_, ny_, nz_}) {
            if (v <= 1) continue;
            if (fft_size_supported(v)) {
                upload_tw(v);
                continue;
            }
            bool have = false;
            for (auto& bt : blue_) have |= (bt.n == v);
            if (have) continue;
            const int m = bluestein_length(v);
            upload_tw(m);
            for (int dir : {-1, +1}) {
                std::vector<cpx<T>> ch, fl;
                bluestein_tables<T>(v, dir, m, ch, fl);
                Blue bt;
                bt.n = v;
                bt.dir = dir;
                bt.m = m;
                bt.chirp.alloc(v);
                bt.filt.alloc(m);
                HIPCHK(hipMemcpy(bt.chirp.get(), ch.data(), v * sizeof(cpx<T>), hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(bt.filt.get(), fl.data(), m * sizeof(cpx<T>), hipMemcpyHostToDevice));
                blue_.push_back(std::move(bt));
            }
        }
    }
    const cpx<T>* twid(int n) const {
        for (auto& p : tw_)
            if (p.first == n) return p.second.get();
        return nullptr;
    }
    void pass(const Seg& s, int n, int dir) {
        if (!fft_size_supported(n)) {
            // no compiled schedule and longer than the brute-force limit: Bluestein
            for (auto& bt : blue_) {
                if (bt.n != n || bt.dir != dir) continue;
                BluesteinArgs<T> b;
                b.p.tw = twid(bt.m);
                b.p.narr = 1;
                b.p.arr[0][0] = d_.get();
                b.p.seg[0] = s;
                b.p.nseg = 1;
                b.p.TW = choose_tw(bt.m, 1, sizeof(cpx<T>), s.sb == 1 ? 8 : 1);
                b.m = bt.m;
                b.chirp = bt.chirp.get();
                b.filt = bt.filt.get();
                HIPCHK(launch_bluestein<T>(b, n, st_));
                return;
            }
            throw Error(GPF_EINVAL, "no Bluestein plan for length " + std::to_string(n));
        }
        PencilArgs<T> a;
        for (auto& p : tw_)
            if (p.first == n) a.tw = p.second.get();
        a.narr = 1;
        a.arr[0][0] = d_.get();
        a.seg[0] = s;
        a.nseg = 1;
        a.TW = choose_tw(n, 1, sizeof(cpx<T>), s.sb == 1 ? 8 : 1);
        HIPCHK(launch_pencils<T>(a, n, dir, st_));
    }
    int ndim_;
    int nx_, ny_, nz_;
    long n_;
    hipStream_t st_ = nullptr;
    Buf<cpx<T>> d_;
    std::vector<std::pair<int, Buf<cpx<T>>>> tw_;
    struct Blue {
        int n = 0, dir = 0, m = 0;
        Buf<cpx<T>> chirp, filt;
    };
    std::vector<Blue> blue_;
};

}  // namespace gpf

// ===========================================================================
// C ABI
// ===========================================================================
using namespace gpf;

struct gpf_model {
    Model m;
    bool initialised = false;
};

struct gpf_solver {
    std::unique_ptr<EngineBase> eng;
    int stepper = GPF_EULER;
    double dt = 0.0;
    long current_step = 0;
    int device = 0;
};

struct gpf_fft {
    std::unique_ptr<FFTBase> impl;
    int device = 0;
};

struct gpf_batch {
    std::unique_ptr<BatchBase> impl;
    int device = 0;
};

template <typename F>
static int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return GPF_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const ParseError& e) {
        g_last_error = e.what();
        return GPF_EPARSE;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return GPF_EINVAL;
    }
}

static void set_device(int dev) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) throw Error(GPF_EDEVICE, "no HIP device available");
    if (dev < 0 || dev >= n) throw Error(GPF_EINVAL, "device ordinal out of range");
    HIPCHK(hipSetDevice(dev));
}

extern "C" {

const char* gpf_last_error(void) { return g_last_error.c_str(); }

int gpf_model_create(gpf_model** out) {
    return guarded([&] {
        if (!out) throw Error(GPF_EINVAL, "null output");
        *out = new gpf_model();
    });
}

void gpf_model_destroy(gpf_model* m) { delete m; }

int gpf_model_add_field(gpf_model* m, const char* name, const double* data, long n) {
    return guarded([&] {
        if (!m || !name || n <= 0) throw Error(GPF_EINVAL, "bad argument");
        std::vector<std::complex<double>> v(n);
        if (data)
            for (long i = 0; i < n; ++i) v[i] = {data[2 * i], data[2 * i + 1]};
        m->m.add_field(name, std::move(v));
    });
}

int gpf_model_add_scalar(gpf_model* m, const char* name, double re, double im) {
    return guarded([&] {
        if (!m || !name) throw Error(GPF_EINVAL, "bad argument");
        m->m.add_scalar(name, {re, im});
    });
}

int gpf_model_add_equation(gpf_model* m, const char* eq) {
    return guarded([&] {
        if (!m || !eq) throw Error(GPF_EINVAL, "bad argument");
        m->m.add_equation(eq);
    });
}

int gpf_model_register_square_gradient(gpf_model* m, const char* term, const char* field, double factor,
                                       int implicit) {
    return guarded([&] {
        if (!m || !term || !field) throw Error(GPF_EINVAL, "bad argument");
        UserTerm u;
        u.kind = kSquaredGradient;
        u.field = field;
        u.factor = factor;
        if (implicit) m->m.register_implicit_term(term, u);
        else m->m.register_explicit_term(term, u);
    });
}

int gpf_model_register_homogeneous_modulus(gpf_model* m, const char* term, const char* field, const double* C,
                                           const double* misfit) {
    return guarded([&] {
        if (!m || !term || !field || !C || !misfit) throw Error(GPF_EINVAL, "bad argument");
        UserTerm u;
        u.kind = kHomogeneousModulus;
        u.field = field;
        std::memcpy(u.C, C, sizeof(u.C));
        std::memcpy(u.misfit, misfit, sizeof(u.misfit));
        m->m.register_explicit_term(term, u);
    });
}

int gpf_model_register_volume_conserving(gpf_model* m, const char* term, const char* field, const char* indicator,
                                         double dt, long num_nodes) {
    return guarded([&] {
        if (!m || !term || !field || !indicator) throw Error(GPF_EINVAL, "bad argument");
        UserTerm u;
        u.kind = kVolumeConserving;
        u.field = field;
        u.indicator = indicator;
        u.dt = dt;
        u.num_nodes = num_nodes;
        m->m.register_explicit_term(term, u);
    });
}

int gpf_model_register_spectral_viscosity(gpf_model* m, const char* term, double eps, double threshold, int power) {
    return guarded([&] {
        if (!m || !term) throw Error(GPF_EINVAL, "bad argument");
        UserTerm u;
        u.kind = kSpectralViscosity;
        u.eps = eps;
        u.threshold = threshold;
        u.power = power;
        m->m.register_implicit_term(term, u);
    });
}

int gpf_model_register_tensorial_hessian(gpf_model* m, const char* term, const char* field, const double* K,
                                         int nk) {
    return guarded([&] {
        if (!m || !term || !K || (nk != 4 && nk != 9)) throw Error(GPF_EINVAL, "bad argument");
        UserTerm u;
        u.kind = kTensorialHessian;
        u.field = field ? field : "";
        u.hess_len = nk;
        for (int k = 0; k < nk; ++k) u.hess[k] = K[k];
        m->m.register_implicit_term(term, u);
    });
}

int gpf_model_register_white_noise(gpf_model* m, const char* name, double strength, unsigned long long seed) {
    return guarded([&] {
        if (!m || !name || !(strength >= 0.0)) throw Error(GPF_EINVAL, "bad argument");
        m->m.register_white_noise(name, strength, seed);
    });
}

int gpf_model_register_conservative_noise(gpf_model* m, const char* term, double strength, int dim,
                                          unsigned long long seed) {
    return guarded([&] {
        if (!m || !term || !(strength >= 0.0)) throw Error(GPF_EINVAL, "bad argument");
        m->m.register_conservative_noise(term, strength, dim, seed);
    });
}

int gpf_model_register_polynomial(gpf_model* m, const char* name, int nterms, const double* coefs,
                                  const double* powers, int nfields) {
    return guarded([&] {
        if (!m || !name || nterms < 0 || (nterms && (!coefs || !powers))) throw Error(GPF_EINVAL, "bad argument");
        if (nfields != (int)m->m.field_names.size())
            throw Error(GPF_EINVAL, "powers must list one exponent per field added so far");
        std::vector<std::pair<double, std::vector<double>>> t;
        for (int i = 0; i < nterms; ++i)
            t.push_back({coefs[i], std::vector<double>(powers + (size_t)i * nfields, powers + (size_t)(i + 1) * nfields)});
        m->m.register_polynomial(name, t);
    });
}

int gpf_model_init(gpf_model* m) {
    return guarded([&] {
        if (!m) throw Error(GPF_EINVAL, "bad argument");
        m->m.init();
        m->initialised = true;
    });
}

int gpf_model_num_bricks(const gpf_model* m, int* nfields, int* nderived) {
    return guarded([&] {
        if (!m) throw Error(GPF_EINVAL, "bad argument");
        if (nfields) *nfields = (int)m->m.field_names.size();
        if (nderived) *nderived = (int)m->m.derived.size();
    });
}

int gpf_model_brick_name(const gpf_model* m, int idx, char* buf, int buflen) {
    return guarded([&] {
        if (!m || !buf || buflen <= 0) throw Error(GPF_EINVAL, "bad argument");
        auto names = m->m.all_field_names();
        if (idx < 0 || idx >= (int)names.size()) throw Error(GPF_EINVAL, "brick index out of range");
        std::strncpy(buf, names[idx].c_str(), buflen - 1);
        buf[buflen - 1] = 0;
    });
}

int gpf_model_num_terms(const gpf_model* m, int eq, int* nterms, int* ndenum) {
    return guarded([&] {
        if (!m || !m->initialised) throw Error(GPF_ESTATE, "model is not initialised");
        if (eq < 0 || eq >= (int)m->m.rhs.size()) throw Error(GPF_EINVAL, "equation index out of range");
        if (nterms) *nterms = (int)m->m.rhs[eq].first.size();
        if (ndenum) *ndenum = (int)m->m.rhs[eq].second.size();
    });
}

int gpf_model_get_term(const gpf_model* m, int eq, int denum, int idx, gpf_term* out) {
    return guarded([&] {
        if (!m || !m->initialised || !out) throw Error(GPF_ESTATE, "model is not initialised");
        if (eq < 0 || eq >= (int)m->m.rhs.size()) throw Error(GPF_EINVAL, "equation index out of range");
        const auto& L = denum ? m->m.rhs[eq].second : m->m.rhs[eq].first;
        if (idx < 0 || idx >= (int)L.size()) throw Error(GPF_EINVAL, "term index out of range");
        const TermDesc& t = L[idx];
        out->kind = t.kind;
        out->coef_re = t.coef.real();
        out->coef_im = t.coef.imag();
        out->brick = t.brick;
        out->has_lap = t.has_lap ? 1 : 0;
        out->lap = t.lap;
        out->user = t.user;
    });
}

void gpf_options_default(gpf_options* o) {
    if (!o) return;
    o->precision = GPF_FP64;
    o->stepper = GPF_EULER;
    o->filter_order = 0;
    o->fused = 1;
    o->device = 0;
    o->autotune = 0;
}

int gpf_solver_create(gpf_model* m, const int* dims, int ndim, double dt, const gpf_options* o, gpf_solver** out) {
    return guarded([&] {
        if (!m || !dims || !out) throw Error(GPF_EINVAL, "bad argument");
        if (!m->initialised) throw Error(GPF_ESTATE, "model is not initialised (call gpf_model_init)");
        gpf_options opt;
        gpf_options_default(&opt);
        if (o) opt = *o;
        if (opt.stepper != GPF_EULER && opt.stepper != GPF_RK4 && opt.stepper != GPF_IMPLICIT_EULER &&
            opt.stepper != GPF_SDD)
            throw Error(GPF_EINVAL, "Unknown stepper scheme");
        set_device(opt.device);
        auto s = std::make_unique<gpf_solver>();
        s->stepper = opt.stepper;
        s->dt = dt;
        s->device = opt.device;
        if (opt.precision == GPF_FP32) s->eng.reset(new Engine<float>(m->m, dims, ndim, dt, opt));
        else if (opt.precision == GPF_FP64) s->eng.reset(new Engine<double>(m->m, dims, ndim, dt, opt));
        else throw Error(GPF_EINVAL, "unknown precision");
        *out = s.release();
    });
}

void gpf_solver_destroy(gpf_solver* s) {
    if (!s) return;
    try {
        set_device(s->device);
    } catch (...) {
    }
    delete s;
}

int gpf_solver_propagate(gpf_solver* s, int nsteps) {
    return guarded([&] {
        if (!s || nsteps < 0) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        if (s->eng->fused()) {
            // the fused class has no term with OnStepFinished state
            s->eng->steps(nsteps);
            if (s->stepper == GPF_EULER) s->current_step += nsteps;  // Euler.Step increments (euler.go:46)
            return;
        }
        for (int i = 0; i < nsteps; ++i) {
            s->eng->step();
            // Euler.Step and ImplicitEuler.Step increment (euler.go:46, implicitEuler.go:206);
            // RK4.Step does not (rk4.go:29-74)
            if (s->stepper != GPF_RK4) s->current_step++;
            s->eng->on_step_finished();
        }
    });
}

int gpf_solver_prepare(gpf_solver* s, int nsteps, double* ms) {
    return guarded([&] {
        if (!s || nsteps < 0) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        const double t = s->eng->prepare(nsteps);
        if (ms) *ms = t;
    });
}

int gpf_solver_stepper_propagate(gpf_solver* s, int nsteps) {
    return guarded([&] {
        if (!s || nsteps < 0) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->steps(nsteps);
        s->current_step += nsteps;  // Euler.Step / RK4.Propagate increment
    });
}

int gpf_solver_adjoint(gpf_solver* s, int nsteps, int checkpoint_every, const double* target, long n,
                       double* loss, double* grad_field, int nscalars, const char* const* scalar_names,
                       double* grad_scalars) {
    return guarded([&] {
        if (!s || !target || !loss) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        if (s->stepper != GPF_EULER) throw Error(GPF_EINVAL, "the adjoint covers the Euler stepper");
        if (n != s->eng->size()) throw Error(GPF_EINVAL, "buffer length does not match the grid");
        s->eng->adjoint(nsteps, checkpoint_every, target, loss, grad_field, nscalars, scalar_names, grad_scalars);
        s->current_step += nsteps;
    });
}

int gpf_solver_get_time(gpf_solver* s, double* t) {
    return guarded([&] {
        if (!s || !t) throw Error(GPF_EINVAL, "bad argument");
        *t = (double)s->current_step * s->dt;
    });
}

int gpf_solver_get_field(gpf_solver* s, int idx, double* out, long n) {
    return guarded([&] {
        if (!s || !out) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        if (n != s->eng->size()) throw Error(GPF_EINVAL, "buffer length does not match the grid");
        s->eng->get_field(idx, out);
    });
}

int gpf_solver_set_field(gpf_solver* s, int idx, const double* in, long n) {
    return guarded([&] {
        if (!s || !in) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        if (n != s->eng->size()) throw Error(GPF_EINVAL, "buffer length does not match the grid");
        s->eng->set_field(idx, in);
    });
}

int gpf_solver_energy(gpf_solver* s, int idx, int npoly, const double* poly, double kappa, double* energy,
                      double* mass) {
    return guarded([&] {
        if (!s || !energy || !mass || (npoly && !poly)) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->energy(idx, npoly, poly, kappa, energy, mass);
    });
}

static void to_reductions(const double* r, gpf_reductions* out) {
    out->sum = r[kRedSum];
    out->sum_sq = r[kRedSumSq];
    out->min = r[kRedMin];
    out->max = r[kRedMax];
    out->nonfinite = (long)r[kRedNonFinite];
}

int gpf_solver_reduce(gpf_solver* s, int idx, gpf_reductions* out) {
    return guarded([&] {
        if (!s || !out) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        double r[kRedOut];
        s->eng->reduce(idx, r);
        to_reductions(r, out);
    });
}

int gpf_reduce(const void* data, long n, int precision, int device, void* stream, gpf_reductions* out) {
    return guarded([&] {
        if ((!data && n > 0) || n < 0 || !out) throw Error(GPF_EINVAL, "bad argument");
        if (precision != GPF_FP32 && precision != GPF_FP64) throw Error(GPF_EINVAL, "unknown precision");
        set_device(device);
        const hipStream_t st = static_cast<hipStream_t>(stream);
        Buf<double> part(kRedOut * reduce_blocks(n)), res(kRedOut);
        if (precision == GPF_FP32)
            HIPCHK(launch_field_reduce<float>(static_cast<const float*>(data), n, nullptr, part.get(), res.get(), st));
        else
            HIPCHK(launch_field_reduce<double>(static_cast<const double*>(data), n, nullptr, part.get(), res.get(), st));
        double r[kRedOut];
        HIPCHK(hipMemcpyAsync(r, res.get(), sizeof(r), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        to_reductions(r, out);
    });
}

int gpf_reduce_noalloc(const void* data, long n, int precision, int device, void* stream, double* result) {
    return guarded([&] {
        if ((!data && n > 0) || n < 0 || !result) throw Error(GPF_EINVAL, "bad argument");
        if (precision != GPF_FP32 && precision != GPF_FP64) throw Error(GPF_EINVAL, "unknown precision");
        if (n > kReduceStaticMaxN)
            throw Error(GPF_EINVAL, "gpf_reduce_noalloc: n = " + std::to_string(n) + " exceeds its maximum of " +
                                        std::to_string(kReduceStaticMaxN) + " values (use gpf_reduce)");
        set_device(device);
        const hipStream_t st = static_cast<hipStream_t>(stream);
        if (precision == GPF_FP32) field_reduce_static<float>(static_cast<const float*>(data), n, result, st);
        else field_reduce_static<double>(static_cast<const double*>(data), n, result, st);
    });
}

int gpf_solver_monitor_energy(gpf_solver* s, int npoly, const double* poly, double kappa) {
    return guarded([&] {
        if (!s || (npoly && !poly)) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->monitor(npoly, poly, kappa);
    });
}

int gpf_solver_energy_series(gpf_solver* s, double* energy, double* mass, long cap, long* n) {
    return guarded([&] {
        if (!s || !n) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        *n = s->eng->series(energy, mass, cap);
    });
}

int gpf_solver_synchronize(gpf_solver* s) {
    return guarded([&] {
        if (!s) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        HIPCHK(hipStreamSynchronize(s->eng->stream()));
        s->eng->after_sync();
    });
}

int gpf_solver_set_persistent(gpf_solver* s, int mode, int threads) {
    return guarded([&] {
        if (!s) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->set_persistent(mode, threads);
    });
}

int gpf_solver_persistent_steps(gpf_solver* s, long* n) {
    return guarded([&] {
        if (!s || !n) throw Error(GPF_EINVAL, "bad argument");
        *n = s->eng->persistent_steps();
    });
}

int gpf_solver_is_fused(const gpf_solver* s) { return (s && s->eng && s->eng->fused()) ? 1 : 0; }

int gpf_solver_enable_timing(gpf_solver* s, int on) {
    return guarded([&] {
        if (!s) throw Error(GPF_EINVAL, "bad argument");
        s->eng->enable_timing(on != 0);
    });
}

int gpf_solver_pass_times(gpf_solver* s, int maxpass, double* ms, long* launches, char* names, int namelen) {
    int np = 0;
    int rc = guarded([&] {
        if (!s) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        np = s->eng->pass_times(maxpass, ms, launches, names, namelen);
    });
    return rc ? -rc : np;
}

int gpf_device_copy_bandwidth(long bytes, int reps, int device, double* ms_per_copy) {
    return guarded([&] {
        if (bytes < 4 || reps < 1 || !ms_per_copy) throw Error(GPF_EINVAL, "bad argument");
        set_device(device);
        const long n = (bytes / (long)sizeof(float)) & ~3L;
        Buf<float> a(n), b(n);
        hipStream_t st;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipMemsetAsync(a.get(), 0, n * sizeof(float), st));
        HIPCHK(launch_copy_bw(a.get(), b.get(), n, st));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) HIPCHK(launch_copy_bw(a.get(), b.get(), n, st));
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_copy = ms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(st);
    });
}

int gpf_device_bandwidth(long bytes, int reps, int kind, int device, double* ms_per_pass) {
    return guarded([&] {
        if (bytes < (1L << 20) || reps < 1 || kind < 0 || kind > 2 || !ms_per_pass)
            throw Error(GPF_EINVAL, "bad argument");
        set_device(device);
        const long unit = 4L * 256 * 4;  // floats per grid-stride step of one thread block row
        const long n = (bytes / (long)sizeof(float)) / unit * unit;
        Buf<float> a(n), b(n);
        hipStream_t st;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipMemsetAsync(a.get(), 0, n * sizeof(float), st));
        HIPCHK(hipMemsetAsync(b.get(), 0, n * sizeof(float), st));
        HIPCHK(launch_bw(kind, a.get(), b.get(), n, st));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) HIPCHK(launch_bw(kind, a.get(), b.get(), n, st));
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_pass = ms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(st);
    });
}

int gpf_solver_set_tuning(gpf_solver* s, const int* tws, int n) {
    return guarded([&] {
        if (!s || !tws) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->set_tuning(tws, n);
    });
}

int gpf_solver_tuning(gpf_solver* s, int* tws, int maxn) {
    if (!s || !s->eng || !tws) return -GPF_EINVAL;
    return s->eng->tuning(tws, maxn);
}

void* gpf_solver_stream(gpf_solver* s) { return (s && s->eng) ? (void*)s->eng->stream() : nullptr; }

int gpf_solver_set_newton_krylov(gpf_solver* s, int maxiter, double step_size, double tol, int stencil,
                                 int restart) {
    return guarded([&] {
        if (!s) throw Error(GPF_EINVAL, "bad argument");
        if (s->stepper != GPF_IMPLICIT_EULER) throw Error(GPF_EINVAL, "the solver's stepper is not ImplicitEuler");
        s->eng->set_newton_krylov(maxiter, step_size, tol, stencil, restart);
    });
}

int gpf_solver_newton_stats(gpf_solver* s, int* newton_iters, int* krylov_iters, int* converged,
                            double* residual) {
    return guarded([&] {
        if (!s) throw Error(GPF_EINVAL, "bad argument");
        if (s->stepper != GPF_IMPLICIT_EULER) throw Error(GPF_EINVAL, "the solver's stepper is not ImplicitEuler");
        s->eng->newton_stats(newton_iters, krylov_iters, converged, residual);
    });
}

int gpf_solver_sdd_init(gpf_solver* s, const double* init, const double* final, long n) {
    return guarded([&] {
        if (!s || !init || !final) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->sdd_init(init, final, n);
    });
}

int gpf_solver_sdd_set_orientation(gpf_solver* s, const double* orient, long n) {
    return guarded([&] {
        if (!s || !orient) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->sdd_set_orientation(orient, n);
    });
}

int gpf_solver_sdd_set_params(gpf_solver* s, double alpha, double tau_orientation, double tau_length,
                              double init_length, double min_length) {
    return guarded([&] {
        if (!s) throw Error(GPF_EINVAL, "bad argument");
        s->eng->sdd_set_params(alpha, tau_orientation, tau_length, init_length, min_length);
    });
}

int gpf_solver_sdd_monitor(gpf_solver* s, double* out) {
    return guarded([&] {
        if (!s || !out) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->sdd_monitor(out);
    });
}

int gpf_solver_sdd_orientation(gpf_solver* s, double* out, long n) {
    return guarded([&] {
        if (!s || !out) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->sdd_orientation(out, n);
    });
}

int gpf_solver_sdd_dimer_length(gpf_solver* s, double t, double* length) {
    return guarded([&] {
        if (!s || !length) throw Error(GPF_EINVAL, "bad argument");
        *length = s->eng->sdd_dimer_length(t);
    });
}

int gpf_solver_save_fields(gpf_solver* s, const char* prefix, int epoch, int format) {
    return guarded([&] {
        if (!s || !prefix) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->save_fields(prefix, epoch, format);
    });
}

int gpf_solver_flush_io(gpf_solver* s) {
    return guarded([&] {
        if (!s) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->flush_io();
    });
}

int gpf_solver_save_csv(gpf_solver* s, const char* prefix, int epoch) {
    return guarded([&] {
        if (!s || !prefix) throw Error(GPF_EINVAL, "bad argument");
        set_device(s->device);
        s->eng->save_csv(prefix, epoch);
    });
}

int gpf_save_float64(const char* path, const double* data, long n) {
    return guarded([&] {
        if (!path || (n > 0 && !data) || n < 0) throw Error(GPF_EINVAL, "bad argument");
        save_float64(path, data, n);
    });
}

int gpf_load_float64(const char* path, double* out, long cap, long* n) {
    return guarded([&] {
        if (!path || !n) throw Error(GPF_EINVAL, "bad argument");
        const std::vector<double> v = load_float64(path);
        *n = (long)v.size();
        if (out) std::memcpy(out, v.data(), std::min((long)v.size(), cap) * sizeof(double));
    });
}

static std::vector<std::string> name_list(const char* const* names, int n) {
    if (n < 0 || (n > 0 && !names)) throw Error(GPF_EINVAL, "bad argument");
    std::vector<std::string> out;
    for (int i = 0; i < n; ++i) {
        if (!names[i]) throw Error(GPF_EINVAL, "bad argument");
        out.emplace_back(names[i]);
    }
    return out;
}

int gpf_create_xdmf(const char* const* fields, int nfields, const char* prefix, int num, const int* dims, int ndim,
                    const char* indent, char* buf, long cap, long* len) {
    return guarded([&] {
        if (!prefix || !dims || !len) throw Error(GPF_EINVAL, "bad argument");
        const std::string t = xdmf_text_indent(name_list(fields, nfields), prefix, num,
                                               std::vector<int>(dims, dims + ndim), indent ? indent : "    ");
        *len = (long)t.size();
        if (buf && cap > 0) {
            const long m = std::min((long)t.size(), cap - 1);
            std::memcpy(buf, t.data(), m);
            buf[m] = '\0';
        }
    });
}

int gpf_write_xdmf(const char* path, const char* const* fields, int nfields, const char* prefix, int num,
                   const int* dims, int ndim) {
    return guarded([&] {
        if (!path || !prefix || !dims) throw Error(GPF_EINVAL, "bad argument");
        const std::string t = xdmf_text(name_list(fields, nfields), prefix, num, std::vector<int>(dims, dims + ndim));
        FILE* f = std::fopen(path, "wb");
        if (!f) throw Error(GPF_EINVAL, std::string("open ") + path + ": cannot create file");
        std::fwrite(t.data(), 1, t.size(), f);
        std::fclose(f);
    });
}

int gpf_save_csv(const char* path, const char* const* names, int nfields, const double* values, long n,
                 const int* dims, int ndim) {
    return guarded([&] {
        if (!path || (nfields > 0 && !values) || n < 0) throw Error(GPF_EINVAL, "bad argument");
        std::vector<std::vector<double>> v(nfields);
        for (int i = 0; i < nfields; ++i) v[i].assign(values + (long)i * n, values + (long)(i + 1) * n);
        write_csv(path, name_list(names, nfields), v, dims ? std::vector<int>(dims, dims + ndim) : std::vector<int>());
    });
}

int gpf_solver_set_filter(gpf_solver* s, const double* table, long n) {
    return guarded([&] {
        if (!s) throw Error(GPF_EINVAL, "bad argument");
        if (table && n != 1000) throw Error(GPF_EINVAL, "a modal filter table has 1000 entries");
        set_device(s->device);
        s->eng->set_filter(table);
    });
}

int gpf_batch_create(gpf_model* m, int nclasses, const int* shapes, const int* counts, double dt,
                     const gpf_options* o, gpf_batch** out) {
    return guarded([&] {
        if (!m || !shapes || !counts || !out || nclasses < 1) throw Error(GPF_EINVAL, "bad argument");
        if (!m->initialised) throw Error(GPF_ESTATE, "model is not initialised (call gpf_model_init)");
        gpf_options opt;
        gpf_options_default(&opt);
        if (o) opt = *o;
        set_device(opt.device);
        auto b = std::make_unique<gpf_batch>();
        b->device = opt.device;
        b->impl.reset(make_batch(m->m, nclasses, shapes, counts, dt, opt));
        *out = b.release();
    });
}

void gpf_batch_destroy(gpf_batch* b) {
    if (!b) return;
    try {
        set_device(b->device);
    } catch (...) {
    }
    delete b;
}

int gpf_batch_num_sims(const gpf_batch* b) { return (b && b->impl) ? b->impl->nsims() : -1; }

int gpf_batch_set_scalar(gpf_batch* b, int sim, const char* name, double value) {
    return guarded([&] {
        if (!b || !name) throw Error(GPF_EINVAL, "bad argument");
        b->impl->set_scalar(sim, name, value);
    });
}

int gpf_batch_set_field(gpf_batch* b, int sim, const double* data, long n) {
    return guarded([&] {
        if (!b || !data) throw Error(GPF_EINVAL, "bad argument");
        set_device(b->device);
        b->impl->set_field(sim, data, n);
    });
}

int gpf_batch_get_field(gpf_batch* b, int sim, double* out, long n) {
    return guarded([&] {
        if (!b || !out) throw Error(GPF_EINVAL, "bad argument");
        set_device(b->device);
        b->impl->get_field(sim, out, n);
    });
}

int gpf_batch_propagate(gpf_batch* b, int nsteps) {
    return guarded([&] {
        if (!b || nsteps < 0) throw Error(GPF_EINVAL, "bad argument");
        set_device(b->device);
        b->impl->propagate(nsteps);
    });
}

int gpf_batch_synchronize(gpf_batch* b) {
    return guarded([&] {
        if (!b) throw Error(GPF_EINVAL, "bad argument");
        set_device(b->device);
        b->impl->synchronize();
    });
}

int gpf_batch_set_layout(gpf_batch* b, const int* v, int n) {
    return guarded([&] {
        if (!b || !v) throw Error(GPF_EINVAL, "bad argument");
        set_device(b->device);
        b->impl->set_layout(v, n);
    });
}

int gpf_batch_tuning(gpf_batch* b, int* tws, int maxn) {
    if (!b || !tws) return -GPF_EINVAL;
    return b->impl->tuning(tws, maxn);
}

int gpf_fft_create(const int* dims, int ndim, int precision, int device, gpf_fft** out) {
    return guarded([&] {
        if (!dims || !out) throw Error(GPF_EINVAL, "bad argument");
        auto f = std::make_unique<gpf_fft>();
        f->device = device;
        if (precision == GPF_FP32) f->impl.reset(new FullFFT<float>(dims, ndim));
        else f->impl.reset(new FullFFT<double>(dims, ndim));
        *out = f.release();
    });
}

void gpf_fft_destroy(gpf_fft* f) {
    if (!f) return;
    try {
        set_device(f->device);
    } catch (...) {
    }
    delete f;
}

int gpf_fft_forward(gpf_fft* f, double* data, long n) {
    return guarded([&] {
        if (!f || !data) throw Error(GPF_EINVAL, "bad argument");
        if (n != f->impl->size()) throw Error(GPF_EINVAL, "buffer length does not match the plan");
        set_device(f->device);
        f->impl->exec(data, -1);
    });
}

int gpf_fft_inverse(gpf_fft* f, double* data, long n) {
    return guarded([&] {
        if (!f || !data) throw Error(GPF_EINVAL, "bad argument");
        if (n != f->impl->size()) throw Error(GPF_EINVAL, "buffer length does not match the plan");
        set_device(f->device);
        f->impl->exec(data, +1);
    });
}

int gpf_fft_freq(const int* dims, int ndim, long i, double* out) {
    return guarded([&] {
        if (!dims || !out || ndim < 2 || ndim > 3) throw Error(GPF_EINVAL, "bad argument");
        const long c = i % dims[1];
        const long r = (i / dims[1]) % dims[0];
        out[1] = (double)c / (double)dims[1];
        out[0] = (double)r / (double)dims[0];
        if (ndim > 2) out[2] = (double)(i / ((long)dims[0] * dims[1])) / (double)dims[2];
        for (int a = 0; a < ndim; ++a)
            if (out[a] > 0.5) out[a] -= 1.0;
    });
}

long gpf_fft_conjugate_node(const int* dims, int ndim, long i) {
    if (!dims || ndim < 2 || ndim > 3) return -1;
    const long c = i % dims[1];
    const long r = (i / dims[1]) % dims[0];
    const long cc = (dims[1] - c) % dims[1];
    const long cr = (dims[0] - r) % dims[0];
    long cd = 0;
    if (ndim == 3) cd = (dims[2] - i / ((long)dims[0] * dims[1])) % dims[2];
    return cd * dims[0] * dims[1] + cr * dims[1] + cc;
}

int gpf_vandeven_table(int order, double* table) {
    return guarded([&] {
        if (!table || order < 1) throw Error(GPF_EINVAL, "bad argument");
        vandeven_table(order, table);
    });
}

double gpf_vandeven_eval(const double* data, double x) {
    const double N = 999.0;
    int idx = (int)(x * N);
    if (idx >= 999) return data[999];
    const double dx = 1.0 / N;
    const double dy = data[idx + 1] - data[idx];
    const double x0 = (double)idx * dx;
    return data[idx] + (x - x0) * dy / dx;
}

}  // extern "C"
