// batch.hip -- many independent single-field simulations of uneven 2D sizes.
//
// A parameter sweep that the reference would run as many pf.Solver instances
// (each sim stored under its own simID in a FieldDB, pf/database.go:67-80)
// runs here as size classes: all sims of one grid shape are stacked into one
// array whose slowest axis is the sim index, so one launch of the fused row
// pass and one of the fused y-pencil update pass advance every sim of a class.
// Per-sim scalars become per-sim coefficient tables (the explicit source
// polynomial, the implicit denominator and the linear explicit terms) and a
// per-sim FP64 mean.  Each class keeps its state as the spectra of its sims and one
// work array (see the single solver's fused step).  Steps are captured in a HIP graph
// and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gpf.h"
#include "batch.h"
#include "fft_core.h"
#include "model.h"
#include "spectral.h"

namespace gpf {

namespace {

#define BCHK(x) HIPCHK(x)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t c) { alloc(c); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t c) {
        if (p) (void)hipFree(p);
        BCHK(hipMalloc(reinterpret_cast<void**>(&p), (c ? c : 1) * sizeof(T)));
        n = c;
    }
};

template <typename T>
std::vector<cpx<T>> twiddles(int n) {
    std::vector<cpx<T>> w(n);
    for (int q = 0; q < n; ++q) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)q / (long double)n;
        w[q].x = (T)cosl(a);
        w[q].y = (T)sinl(a);
    }
    return w;
}

// lines per block of an untuned launch: about 256 threads, and the exchange lines plus a
// twiddle table within 80 KiB of LDS (FP64, 16-byte elements: the largest case)
int pick_tw(int n, int variant) {
    const int T = fft_threads_per_line(n, variant);
    int tw = 1;
    while (tw * 2 * T <= 256) tw *= 2;
    if (tw > 64) tw = 64;
    while (tw > 1 && ((size_t)(n + n / 16 + 1) * (tw + 1) + n) * 16 > 80 * 1024) tw /= 2;
    return tw;
}

// a block of tw lines fits: threads, and LDS for the exchange lines, the row pass's copy of
// the twiddle table and the monitor's partial sums (160 KiB per workgroup on gfx950)
bool tw_ok(int n, int tw, size_t elem, int variant) {
    const size_t lds = ((size_t)(n + n / 16 + 1) * (tw + 1) + n) * elem + (size_t)kMaxBlockThreads * sizeof(double);
    return tw * fft_threads_per_line(n, variant) <= kMaxBlockThreads && lds <= 160 * 1024;
}

}  // namespace

// one size class: nsim sims of ny x nx
template <typename T>
struct BatchClass {
    int ny = 0, nx = 0, nsim = 0, first = 0;
    long nmain = 0, nk = 0, rows = 0;
    DBuf<cpx<T>> S, W, xtw, ytw;  // state spectra; work array (S inverse-transformed along y)
    DBuf<double> mean, gpoly, denom, lin;
    int tw_rows = 0, tw_upd = 0, rows_var = 0;
};

template <typename T>
class BatchEngine : public BatchBase {
public:
    BatchEngine(const Model& m, int ncls, const int* shapes, const int* counts, double dt, const gpf_options& o)
        : base_(m), dt_(dt), autotune_(o.autotune != 0) {
        if (m.field_names.size() != 1) throw BatchError(GPF_EINVAL, "batched sims hold exactly one field");
        if (o.stepper != GPF_EULER) throw BatchError(GPF_EINVAL, "batched sims use the Euler stepper");
        BCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        int first = 0;
        for (int c = 0; c < ncls; ++c) {
            auto cl = std::make_unique<BatchClass<T>>();
            cl->ny = shapes[2 * c];
            cl->nx = shapes[2 * c + 1];
            cl->nsim = counts[c];
            cl->first = first;
            if (cl->nx % 2 || cl->ny % 2 || !fft_size_supported(cl->nx) || !fft_size_supported(cl->ny) || cl->nsim < 1)
                throw BatchError(GPF_EINVAL, "unsupported batch class shape (both sides even, supported lengths)");
            cl->rows = (long)cl->nsim * cl->ny;
            cl->nmain = cl->rows * (cl->nx / 2);
            cl->nk = cl->nmain + cl->rows;
            cl->S.alloc(cl->nk);
            cl->W.alloc(cl->nk);
            cl->mean.alloc(cl->nsim);
            BCHK(hipMemsetAsync(cl->S.p, 0, cl->nk * sizeof(cpx<T>), st_));
            BCHK(hipMemsetAsync(cl->W.p, 0, cl->nk * sizeof(cpx<T>), st_));
            BCHK(hipMemsetAsync(cl->mean.p, 0, cl->nsim * sizeof(double), st_));
            auto wx = twiddles<T>(cl->nx), wy = twiddles<T>(cl->ny);
            cl->xtw.alloc(cl->nx);
            cl->ytw.alloc(cl->ny);
            BCHK(hipMemcpy(cl->xtw.p, wx.data(), wx.size() * sizeof(cpx<T>), hipMemcpyHostToDevice));
            BCHK(hipMemcpy(cl->ytw.p, wy.data(), wy.size() * sizeof(cpx<T>), hipMemcpyHostToDevice));
            first += cl->nsim;
            cls_.push_back(std::move(cl));
        }
        nsims_ = first;
        scalars_.assign(nsims_, m.scalars);
        dirty_ = true;
        BCHK(hipStreamSynchronize(st_));
    }

    ~BatchEngine() override {
        drop_graphs();
        if (st_) {
            (void)hipStreamSynchronize(st_);
            (void)hipStreamDestroy(st_);
        }
    }

    int nsims() const override { return nsims_; }

    void set_scalar(int sim, const std::string& name, double v) override {
        check_sim(sim);
        if (!base_.scalars.count(name)) throw BatchError(GPF_EINVAL, "unknown scalar " + name);
        scalars_[sim][name] = {v, 0.0};
        dirty_ = true;
    }

    void set_field(int sim, const double* in, long n) override {
        BatchClass<T>& c = cls_of(sim);
        const int ls = sim - c.first;
        if (n != (long)c.ny * c.nx) throw BatchError(GPF_EINVAL, "buffer length does not match the sim grid");
        double s = 0.0, comp = 0.0;
        for (long j = 0; j < n; ++j) {
            const double v = in[2 * j], t = s + v;
            comp += (std::fabs(s) >= std::fabs(v)) ? ((s - t) + v) : ((v - t) + s);
            s = t;
        }
        const bool split = sizeof(T) == 4;
        const double mean = split ? (s + comp) / (double)n : 0.0;
        std::vector<T> h(n);
        for (long j = 0; j < n; ++j) h[j] = (T)(in[2 * j] - mean);
        DBuf<T> tmp(n);
        BCHK(hipMemcpyAsync(tmp.p, h.data(), n * sizeof(T), hipMemcpyHostToDevice, st_));
        cpx<T>* sm = c.S.p + (long)ls * c.ny * (c.nx / 2);
        cpx<T>* sq = c.S.p + c.nmain + (long)ls * c.ny;
        RowR2CArgs<T> a;
        a.tw = c.xtw.p;
        a.narr = 1;
        a.in[0] = tmp.p;
        a.main[0] = sm;
        a.nyq[0] = sq;
        a.rows = c.ny;
        a.nx = c.nx;
        a.TW = pick_tw(c.nx, 0);
        a.in_scale = (T)(1.0 / ((double)c.nx * c.ny));
        BCHK(launch_rows_r2c<T>(a, st_));
        sim_pencils(c, sm, sq, nullptr, nullptr, -1);
        sim_pencils(c, c.W.p + (sm - c.S.p), c.W.p + (sq - c.S.p), sm, sq, +1);
        BCHK(hipMemcpyAsync(c.mean.p + ls, &mean, sizeof(double), hipMemcpyHostToDevice, st_));
        BCHK(hipStreamSynchronize(st_));
    }

    void get_field(int sim, double* out, long n) override {
        BatchClass<T>& c = cls_of(sim);
        const int ls = sim - c.first;
        if (n != (long)c.ny * c.nx) throw BatchError(GPF_EINVAL, "buffer length does not match the sim grid");
        DBuf<T> tmp(n);
        DBuf<cpx<T>> tk((size_t)c.ny * (c.nx / 2) + c.ny);
        sim_pencils(c, tk.p, tk.p + (long)c.ny * (c.nx / 2), c.S.p + (long)ls * c.ny * (c.nx / 2),
                    c.S.p + c.nmain + (long)ls * c.ny, +1);
        RowC2RArgs<T> a;
        a.tw = c.xtw.p;
        a.narr = 1;
        a.main[0] = tk.p;
        a.nyq[0] = tk.p + (long)c.ny * (c.nx / 2);
        a.out[0] = tmp.p;
        a.rows = c.ny;
        a.nx = c.nx;
        a.TW = pick_tw(c.nx, 0);
        a.scale = T(1);
        BCHK(launch_rows_c2r<T>(a, st_));
        std::vector<T> h(n);
        double mean = 0.0;
        BCHK(hipMemcpyAsync(h.data(), tmp.p, n * sizeof(T), hipMemcpyDeviceToHost, st_));
        BCHK(hipMemcpyAsync(&mean, c.mean.p + ls, sizeof(double), hipMemcpyDeviceToHost, st_));
        BCHK(hipStreamSynchronize(st_));
        if (sizeof(T) != 4) mean = 0.0;
        for (long j = 0; j < n; ++j) {
            out[2 * j] = (double)h[j] + mean;
            out[2 * j + 1] = 0.0;
        }
    }

    void propagate(int nsteps) override {
        if (dirty_) compile();
        if (autotune_ && !tuned_) tune();
        int done = 0;
        // graphs of kGraphSteps steps; the state and work arrays stay in place
        if (nsteps >= kGraphSteps) {
            if (!graph_exec_) capture();
            while (nsteps - done >= kGraphSteps) {
                BCHK(hipGraphLaunch(graph_exec_, st_));
                done += kGraphSteps;
            }
        }
        for (; done < nsteps; ++done) enqueue_step();
    }

    void synchronize() override { BCHK(hipStreamSynchronize(st_)); }

    // per class: rows lines per block, update lines per block, rows schedule; then the
    // threads per block of the grouped launches (0: one launch per class and pass)
    int tuning(int* out, int maxn) override {
        int k = 0;
        for (auto& c : cls_) {
            for (int v : {c->tw_rows, c->tw_upd, c->rows_var}) {
                if (k < maxn) out[k] = v;
                ++k;
            }
        }
        if (k < maxn) out[k] = grouped_ ? gthreads_ : 0;
        ++k;
        return k;
    }

    // a layout in tuning()'s format, applied instead of the tuner
    void set_layout(const int* v, int n) override {
        const int nc = (int)cls_.size();
        if (n != 3 * nc + 1) throw BatchError(GPF_EINVAL, "a batch layout has 3 values per class and 1 more");
        if (dirty_) compile();
        for (int k = 0; k < nc; ++k) {
            BatchClass<T>& c = *cls_[k];
            const int tr = v[3 * k], tu = v[3 * k + 1], var = v[3 * k + 2];
            auto pow2 = [](int x) { return x >= 1 && (x & (x - 1)) == 0; };
            if (!pow2(tr) || !pow2(tu) || (var != 0 && var != 1) || !tw_ok(c.nx, tr, sizeof(cpx<T>), var) ||
                !tw_ok(c.ny, tu, sizeof(cpx<T>), 1))
                throw BatchError(GPF_EINVAL, "batch layout not admissible for class " + std::to_string(k));
        }
        const int th = v[3 * nc];
        if (th != 0 && (!group_ok() || (th != 128 && th != 256 && th != 512)))
            throw BatchError(GPF_EINVAL, "grouped launches need power-of-two sides 64..1024 and 128, 256 or 512 threads");
        BCHK(hipStreamSynchronize(st_));
        for (int k = 0; k < nc; ++k) {
            cls_[k]->tw_rows = v[3 * k];
            cls_[k]->tw_upd = v[3 * k + 1];
            cls_[k]->rows_var = v[3 * k + 2];
        }
        grouped_ = false;
        if (th != 0) {
            if (!build_group(th, own_S(), own_W(), own_mean()))
                throw BatchError(GPF_EINVAL, "grouped launches do not fit these classes at this block size");
            grouped_ = true;
        }
        tuned_ = true;
        drop_graphs();
    }

private:
    static constexpr int kGraphSteps = 10;

    void check_sim(int sim) const {
        if (sim < 0 || sim >= nsims_) throw BatchError(GPF_EINVAL, "sim index out of range");
    }

    BatchClass<T>& cls_of(int sim) {
        check_sim(sim);
        for (auto& c : cls_)
            if (sim >= c->first && sim < c->first + c->nsim) return *c;
        throw BatchError(GPF_EINVAL, "sim index out of range");
    }

    // per-sim coefficients: parse the model with each sim's scalars
    void compile() {
        std::vector<double> gp(kMaxPoly * (size_t)nsims_, 0.0), de(kMaxKTerms * (size_t)nsims_, 0.0),
            li(kMaxKTerms * (size_t)nsims_, 0.0);
        bool first = true;
        dense_ = -1;
        for (int s = 0; s < nsims_; ++s) {
            Model m = base_;
            m.scalars = scalars_[s];
            m.init();
            FusedShape fs;
            std::string why;
            if (!fused_shape(m, &fs, &why)) throw BatchError(GPF_EINVAL, "batched sims need the fused class: " + why);
            if (first) {
                shape_ = fs;
                first = false;
            } else if (!same_structure(shape_, fs)) {
                throw BatchError(GPF_EINVAL, "every sim must have the same term structure");
            }
            for (int q = 0; q < kMaxPoly && q < (int)fs.gpoly.size(); ++q) gp[s * kMaxPoly + q] = fs.gpoly[q];
            if (dense_ < 0) {
                PencilUpdateArgs<T> probe;
                probe.denom = fs.denom;
                probe.lin = fs.lin;
                probe.src_lap = fs.src_lap;
                dense_ = make_dense(probe);
            }
            // per-sim rows: dense coefficients by power, or per-term coefficients
            for (int t = 0; t < fs.denom.n; ++t) {
                if (dense_) de[s * kMaxKTerms + fs.denom.pow[t]] += fs.denom.coef[t];
                else de[s * kMaxKTerms + t] = fs.denom.coef[t];
            }
            for (int t = 0; t < fs.lin.n; ++t) {
                if (dense_) li[s * kMaxKTerms + fs.lin.pow[t]] += fs.lin.coef[t];
                else li[s * kMaxKTerms + t] = fs.lin.coef[t];
            }
        }
        npoly_ = 0;
        for (int s = 0; s < nsims_; ++s)
            for (int q = 0; q < kMaxPoly; ++q)
                if (gp[s * kMaxPoly + q] != 0.0 && q + 1 > npoly_) npoly_ = q + 1;
        for (auto& c : cls_) {
            c->gpoly.alloc((size_t)kMaxPoly * c->nsim);
            c->denom.alloc((size_t)kMaxKTerms * c->nsim);
            c->lin.alloc((size_t)kMaxKTerms * c->nsim);
            BCHK(hipMemcpy(c->gpoly.p, gp.data() + (size_t)kMaxPoly * c->first, (size_t)kMaxPoly * c->nsim * sizeof(double),
                           hipMemcpyHostToDevice));
            BCHK(hipMemcpy(c->denom.p, de.data() + (size_t)kMaxKTerms * c->first,
                           (size_t)kMaxKTerms * c->nsim * sizeof(double), hipMemcpyHostToDevice));
            BCHK(hipMemcpy(c->lin.p, li.data() + (size_t)kMaxKTerms * c->first, (size_t)kMaxKTerms * c->nsim * sizeof(double),
                           hipMemcpyHostToDevice));
        }
        drop_graphs();
        dirty_ = false;
        // the coefficient tables were reallocated: the grouped launches' argument arrays point
        // at the new ones
        if (grouped_ && !build_group(gthreads_, own_S(), own_W(), own_mean()))
            throw BatchError(GPF_EINVAL, "grouped launch layout no longer fits");
    }

    // y transform of one sim's region (main plane and x-Nyquist column): dst = F_y^dir(src),
    // in place when src is null
    void sim_pencils(BatchClass<T>& c, cpx<T>* dm, cpx<T>* dq, const cpx<T>* sm, const cpx<T>* sq, int dir) {
        PencilArgs<T> a;
        a.tw = c.ytw.p;
        a.narr = 1;
        Seg m;
        m.na = 1; m.nb = c.nx / 2; m.sa = 0; m.sb = 1; m.es = c.nx / 2;
        Seg q;
        q.na = 1; q.nb = 1; q.sa = 0; q.sb = 1; q.es = 1;
        a.seg[0] = m;
        a.seg[1] = q;
        a.nseg = 2;
        a.TW = pick_tw(c.ny, 0);
        a.arr[0][0] = dm;
        a.arr[0][1] = dq;
        a.src[0][0] = sm;
        a.src[0][1] = sq;
        BCHK(launch_pencils<T>(a, c.ny, dir, st_));
    }

    RowFusedArgs<T> rows_args(BatchClass<T>& c, cpx<T>* W, double* mean, int tw, int var) {
        RowFusedArgs<T> ra;
        ra.tw = c.xtw.p;
        ra.s_main = W;
        ra.s_nyq = W + c.nmain;
        ra.g_main = W;
        ra.g_nyq = W + c.nmain;
        ra.rows = c.rows;
        ra.nx = c.nx;
        ra.TW = tw > 0 ? tw : pick_tw(c.nx, var);
        ra.variant = var;
        ra.npoly = npoly_;
        ra.mean = (sizeof(T) == 4) ? mean : nullptr;
        ra.rows_per_sim = c.ny;
        ra.gpoly_sim = c.gpoly.p;
        return ra;
    }

    PencilUpdateArgs<T> upd_args(BatchClass<T>& c, cpx<T>* S, cpx<T>* W, double* mean, int tw) {
        PencilUpdateArgs<T> u;
        u.p.tw = c.ytw.p;
        u.p.narr = 2;
        u.p.arr[0][0] = S;
        u.p.arr[0][1] = S + c.nmain;
        u.p.arr[1][0] = W;
        u.p.arr[1][1] = W + c.nmain;
        Seg m;  // y-pencils of every sim: a = sim, b = kx
        m.na = c.nsim; m.nb = c.nx / 2; m.sa = (long)c.ny * (c.nx / 2); m.sb = 1; m.es = c.nx / 2;
        m.ksrc[0] = 1; m.ksrc[1] = 2; m.ksrc[2] = 3; m.sim_src = 0;
        Seg q;  // x-Nyquist column of every sim: contiguous line per sim
        q.na = 1; q.nb = c.nsim; q.sa = 0; q.sb = c.ny; q.es = 1;
        q.ksrc[0] = 3; q.kfix[0] = c.nx / 2; q.ksrc[1] = 2; q.ksrc[2] = 3; q.sim_src = 1;
        u.p.seg[0] = m;
        u.p.seg[1] = q;
        u.p.nseg = 2;
        u.p.TW = tw > 0 ? tw : pick_tw(c.ny, 1);
        u.g.nx = c.nx;
        u.g.ny = c.ny;
        u.g.nz = 1;
        u.g.ndim = 2;
        u.dt = dt_;
        u.scale = 1.0 / ((double)c.nx * c.ny);
        u.denom = shape_.denom;
        u.lin = shape_.lin;
        u.src_lap = shape_.src_lap;
        u.has_src = shape_.has_src ? 1 : 0;
        u.mean = (sizeof(T) == 4) ? mean : nullptr;
        u.dense = dense_;
        u.denom_sim = c.denom.p;
        u.lin_sim = c.lin.p;
        return u;
    }

    void enqueue_step() {
        if (grouped_) {
            if (shape_.has_src) BCHK(launch_rows_fused_group<T>(grows_, gthreads_, grows_smem_, st_));
            BCHK(launch_pencil_update_group<T>(gupd_, gthreads_, gupd_smem_, st_));
            return;
        }
        for (auto& cp : cls_) {
            BatchClass<T>& c = *cp;
            if (shape_.has_src) BCHK(launch_rows_fused<T>(rows_args(c, c.W.p, c.mean.p, c.tw_rows, c.rows_var), st_));
            BCHK(launch_pencil_update<T>(upd_args(c, c.S.p, c.W.p, c.mean.p, c.tw_upd), c.ny, st_));
        }
    }

    // every class has a length the grouped launches are compiled for
    bool group_ok() const {
        if ((int)cls_.size() > kMaxGroup) return false;
        for (auto& c : cls_)
            if (!group_length(c->nx) || !group_length(c->ny)) return false;
        return true;
    }

    // the grouped launches of one step for `threads` threads per block over the arrays S[c],
    // W[c] and means mu[c] of each class; false when a class's lines per block would not fit
    bool build_group(int threads, const std::vector<cpx<T>*>& S, const std::vector<cpx<T>*>& W,
                     const std::vector<double*>& mu) {
        const int nc = (int)cls_.size();
        std::vector<RowFusedArgs<T>> ra(nc);
        std::vector<PencilUpdateArgs<T>> ua(nc);
        BatchGroupArgs<T> gr, gu;
        gr.ncls = gu.ncls = nc;
        size_t sr = 0, su = 0;
        int fr = 0, fu = 0;
        const size_t el = sizeof(cpx<T>);
        for (int k = 0; k < nc; ++k) {
            BatchClass<T>& c = *cls_[k];
            const int tr = threads / fft_threads_per_line(c.nx, c.rows_var);
            const int tu = threads / fft_threads_per_line(c.ny, 1);
            if (tr < 1 || tu < 1 || tr * fft_threads_per_line(c.nx, c.rows_var) != threads ||
                tu * fft_threads_per_line(c.ny, 1) != threads || !tw_ok(c.nx, tr, el, c.rows_var) || !tw_ok(c.ny, tu, el, 1))
                return false;
            ra[k] = rows_args(c, W[k], mu[k], tr, c.rows_var);
            ua[k] = upd_args(c, S[k], W[k], mu[k], tu);
            ua[k].p.n = c.ny;
            ua[k].p.lpt = 1;
            const long br = rows_fused_blocks(c.rows, tr);
            const long bu = pencil_blocks(ua[k].p.seg, ua[k].p.nseg, tu, &ua[k].p.tiles0);
            gr.first[k] = fr;
            gr.blocks[k] = (int)br;
            gu.first[k] = fu;
            gu.blocks[k] = (int)bu;
            fr += (int)((br + 7) / 8 * 8);
            fu += (int)((bu + 7) / 8 * 8);
            sr = std::max(sr, ((size_t)(c.nx + c.nx / 16 + 1) * (tr + 1) + c.nx) * el);
            su = std::max(su, (size_t)(c.ny + c.ny / 16 + 1) * (tu + 1) * el);
        }
        gr.first[nc] = fr;
        gu.first[nc] = fu;
        growsbuf_.alloc(nc);
        gupdbuf_.alloc(nc);
        BCHK(hipMemcpyAsync(growsbuf_.p, ra.data(), nc * sizeof(RowFusedArgs<T>), hipMemcpyHostToDevice, st_));
        BCHK(hipMemcpyAsync(gupdbuf_.p, ua.data(), nc * sizeof(PencilUpdateArgs<T>), hipMemcpyHostToDevice, st_));
        BCHK(hipStreamSynchronize(st_));
        gr.rows = growsbuf_.p;
        gu.upd = gupdbuf_.p;
        grows_ = gr;
        gupd_ = gu;
        grows_smem_ = sr;
        gupd_smem_ = su;
        gthreads_ = threads;
        return true;
    }

    std::vector<cpx<T>*> own_S() const {
        std::vector<cpx<T>*> v;
        for (auto& c : cls_) v.push_back(c->S.p);
        return v;
    }
    std::vector<cpx<T>*> own_W() const {
        std::vector<cpx<T>*> v;
        for (auto& c : cls_) v.push_back(c->W.p);
        return v;
    }
    std::vector<double*> own_mean() const {
        std::vector<double*> v;
        for (auto& c : cls_) v.push_back(c->mean.p);
        return v;
    }

    void capture() {
        hipGraph_t g = nullptr;
        BCHK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < kGraphSteps; ++s) enqueue_step();
        BCHK(hipStreamEndCapture(st_, &g));
        BCHK(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
    }

    // batch layout autotuner: lines per block of both passes of every class, timed on scratch copies
    void tune() {
        hipEvent_t e0, e1;
        BCHK(hipEventCreate(&e0));
        BCHK(hipEventCreate(&e1));
        const int cand[] = {4, 8, 16, 32, 64};
        for (auto& cp : cls_) {
            BatchClass<T>& c = *cp;
            DBuf<cpx<T>> S(c.nk), W(c.nk);
            DBuf<double> mu(c.nsim);
            BCHK(hipMemcpyAsync(S.p, c.S.p, c.nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
            BCHK(hipMemcpyAsync(W.p, c.W.p, c.nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
            BCHK(hipMemcpyAsync(mu.p, c.mean.p, c.nsim * sizeof(double), hipMemcpyDeviceToDevice, st_));
            auto time_it = [&](auto&& launch) {
                launch();
                BCHK(hipEventRecord(e0, st_));
                for (int r = 0; r < 3; ++r) launch();
                BCHK(hipEventRecord(e1, st_));
                BCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                BCHK(hipEventElapsedTime(&ms, e0, e1));
                return ms;
            };
            float best = 1e30f;
            for (int var = 0; var < 2; ++var)
                for (int tw : cand) {
                    if (!tw_ok(c.nx, tw, sizeof(cpx<T>), var)) continue;
                    const float t = time_it([&] { BCHK(launch_rows_fused<T>(rows_args(c, W.p, mu.p, tw, var), st_)); });
                    if (t < best) { best = t; c.tw_rows = tw; c.rows_var = var; }
                }
            best = 1e30f;
            for (int tw : cand) {
                if (!tw_ok(c.ny, tw, sizeof(cpx<T>), 1)) continue;
                const float t = time_it([&] { BCHK(launch_pencil_update<T>(upd_args(c, S.p, W.p, mu.p, tw), c.ny, st_)); });
                if (t < best) { best = t; c.tw_upd = tw; }
            }
            BCHK(hipStreamSynchronize(st_));
        }
        // packing: per-class launches (2 per class and step) against grouped launches (2 per
        // step) of 128, 256 or 512 threads per block, timed over a whole step on scratch copies
        grouped_ = false;
        if (group_ok()) {
            std::vector<DBuf<cpx<T>>> Sb(cls_.size()), Wb(cls_.size());
            std::vector<DBuf<double>> mb(cls_.size());
            std::vector<cpx<T>*> Sp, Wp;
            std::vector<double*> mp;
            for (size_t k = 0; k < cls_.size(); ++k) {
                BatchClass<T>& c = *cls_[k];
                Sb[k].alloc(c.nk);
                Wb[k].alloc(c.nk);
                mb[k].alloc(c.nsim);
                BCHK(hipMemcpyAsync(Sb[k].p, c.S.p, c.nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
                BCHK(hipMemcpyAsync(Wb[k].p, c.W.p, c.nk * sizeof(cpx<T>), hipMemcpyDeviceToDevice, st_));
                BCHK(hipMemcpyAsync(mb[k].p, c.mean.p, c.nsim * sizeof(double), hipMemcpyDeviceToDevice, st_));
                Sp.push_back(Sb[k].p);
                Wp.push_back(Wb[k].p);
                mp.push_back(mb[k].p);
            }
            auto time_step = [&](bool grouped) {
                auto one = [&] {
                    if (grouped) {
                        if (shape_.has_src) BCHK(launch_rows_fused_group<T>(grows_, gthreads_, grows_smem_, st_));
                        BCHK(launch_pencil_update_group<T>(gupd_, gthreads_, gupd_smem_, st_));
                    } else {
                        for (size_t k = 0; k < cls_.size(); ++k) {
                            BatchClass<T>& c = *cls_[k];
                            if (shape_.has_src)
                                BCHK(launch_rows_fused<T>(rows_args(c, Wp[k], mp[k], c.tw_rows, c.rows_var), st_));
                            BCHK(launch_pencil_update<T>(upd_args(c, Sp[k], Wp[k], mp[k], c.tw_upd), c.ny, st_));
                        }
                    }
                };
                one();
                BCHK(hipEventRecord(e0, st_));
                for (int r = 0; r < 5; ++r) one();
                BCHK(hipEventRecord(e1, st_));
                BCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                BCHK(hipEventElapsedTime(&ms, e0, e1));
                return ms;
            };
            float best = time_step(false);
            int best_threads = 0;
            for (int th : {128, 256, 512}) {
                if (!build_group(th, Sp, Wp, mp)) continue;
                const float t = time_step(true);
                if (t < best) { best = t; best_threads = th; }
            }
            BCHK(hipStreamSynchronize(st_));
            if (best_threads > 0) {
                if (!build_group(best_threads, own_S(), own_W(), own_mean()))
                    throw BatchError(GPF_EINVAL, "grouped launch layout no longer fits");
                grouped_ = true;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        drop_graphs();
        tuned_ = true;
    }

    Model base_;
    double dt_;
    bool autotune_ = false, tuned_ = false, dirty_ = true;
    hipStream_t st_ = nullptr;
    std::vector<std::unique_ptr<BatchClass<T>>> cls_;
    std::vector<std::map<std::string, std::complex<double>>> scalars_;
    int nsims_ = 0, npoly_ = 0;
    int dense_ = 0;  // dense level of the per-sim multiplier rows (make_dense)
    FusedShape shape_;
    hipGraphExec_t graph_exec_ = nullptr;
    // grouped launches: argument arrays in device memory and the launches' shape
    bool grouped_ = false;
    int gthreads_ = 256;
    DBuf<RowFusedArgs<T>> growsbuf_;
    DBuf<PencilUpdateArgs<T>> gupdbuf_;
    BatchGroupArgs<T> grows_, gupd_;
    size_t grows_smem_ = 0, gupd_smem_ = 0;

    void drop_graphs() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
    }
};

BatchBase* make_batch(const Model& m, int ncls, const int* shapes, const int* counts, double dt, const gpf_options& o) {
    if (o.precision == GPF_FP32) return new BatchEngine<float>(m, ncls, shapes, counts, dt, o);
    return new BatchEngine<double>(m, ncls, shapes, counts, dt, o);
}

}  // namespace gpf
