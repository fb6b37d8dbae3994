// spectral.h -- host-visible launch interface of the HIP spectral kernels.
//
// Data layout in HBM (all arrays device-resident across steps):
//   real-space field      T[nz][ny][nx]                 (nx fastest; nz = 1 in 2D)
//   half spectrum along x split in two arrays:
//     main                cpx<T>[nz][ny][nx/2]           kx = 0 .. nx/2-1
//     nyq                 cpx<T>[nz][ny]                 kx = nx/2 (the x-Nyquist plane)
//   full complex array    cpx<T>[nz][ny][nx]             (FourierTransform drop-in)
// The x index of the layout is the reference's "col", y is "row" and z is
// "depth" (pfutil/indexPositionConversion.go:9-11, FFTW plan n = [D0, D1, D2]).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gpf {

template <typename T>
struct cpx;

constexpr int kMaxArr = 4;
constexpr int kMaxBlockThreads = 512;  // threads per block of every transform kernel
constexpr int kMaxPoly = 9;     // polynomial degree < kMaxPoly
constexpr int kMaxKTerms = 8;   // k-space multiplier terms
constexpr int kDenseDeg = 4;    // dense multiplier polynomials in L up to this degree

// line grid of one pencil segment
struct Seg {
    int na = 0, nb = 0;         // lines: a in [0, na), b in [0, nb); tile = TW consecutive b
    long sa = 0, sb = 0, es = 0;
    int ksrc[3] = {3, 3, 3};    // source of (kx, ky, kz): 0 = a, 1 = b, 2 = p, 3 = fixed
    int kfix[3] = {0, 0, 0};
    int sim_src = -1;           // batched sims: index of the sim from 0 = a, 1 = b; -1 none
};

// k-space multiplier sum  M(k) = sum_t coef_t * L(k)^pow_t,  L = -(2 pi |f|)^2
struct KPoly {
    int n = 0;
    double coef[kMaxKTerms];
    int pow[kMaxKTerms];
};

// Division of 32-bit unsigned numerators by a fixed divisor d >= 1 with one multiply-high
// (Granlund & Montgomery, "Division by invariant integers using multiplication", 1994,
// Thm 4.2 with N = 32): l = ceil(log2 d) and m = ceil(2^(32+l) / d) = 2^32 + mp, so that
// q = floor(m n / 2^(32+l)) = (n + umulhi(mp, n)) >> l for every n < 2^32.
struct FastDiv {
    unsigned mp = 0;
    int l = 0;
    unsigned d = 1;
    static FastDiv make(unsigned d) {
        FastDiv f;
        f.d = d;
        while ((1ull << f.l) < d) ++f.l;
        const unsigned long long m = ((1ull << (32 + f.l)) + d - 1) / d;  // 2^32 <= m < 2^33
        f.mp = (unsigned)(m - (1ull << 32));
        return f;
    }
    __host__ __device__ unsigned div(unsigned n) const {
        const unsigned hi = (unsigned)(((unsigned long long)mp * n) >> 32);  // v_mul_hi_u32
        return (unsigned)(((unsigned long long)n + hi) >> l);
    }
};

// geometry of the transform grid
struct KGrid {
    int nx = 1, ny = 1, nz = 1;
    int ndim = 2;
    // the k-space kernels index the half spectrum with 32-bit numerators (the grid keeps
    // nk < 2^32) divided by nx / 2 and ny
    FastDiv dxh, dy;
    // Freq components (fftWrap.go:57-74) per axis: [ny] row (y) | [nx/2 + 1] column (x) |
    // [nz] depth (z, zeros in 2D); each entry is float64(k) / float64(n), less 1 above 0.5
    const double* ftab = nullptr;
};

template <typename T>
struct PencilArgs {
    const cpx<T>* tw = nullptr;          // twiddle table for the pencil length
    cpx<T>* arr[kMaxArr][2] = {};        // [array][segment] output (and input when src is null)
    const cpx<T>* src[kMaxArr][2] = {};  // optional separate input (out-of-place pass)
    int narr = 0;
    Seg seg[2];
    int nseg = 1;
    int tiles0 = 0;                      // tiles of segment 0 (segment 1 follows)
    int TW = 16;                         // lines per block (power of two)
    int n = 0;                           // pencil length (runtime sizes)
    int lpt = 1;                         // lines per thread (k_pencils: 1 or 2; a tile is TW * lpt lines)
    int variant = 0;                     // transform schedule of k_pencils: 0 plain, 1 low-register
};

// Bluestein transform of pencils of any length n <= 1024 (those without a compiled
// schedule): X_k = w_k sum_j (x_j w_j) conj(w_{k-j}), w_j = exp(DIR pi i j^2 / n), a
// circular convolution of length M (a power of two >= 2n - 1) computed with two M-point
// transforms.  chirp holds w_j (n values), filt the forward M-point transform of the
// zero-padded conj(w) divided by M.
template <typename T>
struct BluesteinArgs {
    PencilArgs<T> p;                     // tiles, segments and the M-point twiddle table (p.tw)
    int m = 0;                           // convolution length M
    const cpx<T>* chirp = nullptr;
    const cpx<T>* filt = nullptr;
};

// fused update pass of the single-field fast path, along the last axis of the step:
//   W (source g, transformed along every other axis) --forward--> G^
//   S <- m1 S + m2 G^ / N            (S: the state spectrum F(c - mean) / N, in place)
//   W <- inverse transform of the new S along this axis (the next step's row-pass input)
template <typename T>
struct PencilUpdateArgs {
    PencilArgs<T> p;                     // arr[0] = state spectrum S (in/out), arr[1] = work W (in/out)
    KGrid g;
    double dt = 0.0;
    double scale = 1.0;                  // 1 / N_total applied to the source
    KPoly denom;                         // implicit part D(k)
    KPoly lin;                           // explicit terms linear in the field E(k)
    int src_lap = 0;                     // LAP power of the explicit source
    int has_src = 0;
    const double* filter = nullptr;      // Vandeven table (1000 entries) or null
    double* mean = nullptr;              // FP64 k = 0 mode: the field mean (in/out), or null when the
                                         // k = 0 coefficient is carried by the arrays themselves
    double* egrad_partial = nullptr;     // per-block FP64 sum of w (2 pi |f|)^2 |FFT(c)|^2 or null
    // dense form (dense = 2 or 4: every LAP power of D, E and the source is at most that
    // degree): D(L) = sum_p dd[p] L^p, E(L) = sum_p dl[p] L^p; 0: the term lists above
    int dense = 0;
    double dd[kDenseDeg + 1] = {};
    double dl[kDenseDeg + 1] = {};
    // batched sims (Seg::sim_src >= 0): per-sim coefficients of denom / lin and mean[sim],
    // rows of kMaxKTerms: indexed by power when dense, else by term (powers from denom / lin)
    const double* denom_sim = nullptr;
    const double* lin_sim = nullptr;
    // tabulated multipliers (the persistent step): s1 = m1 - 1 and s2 = m2 / N of every
    // k-point, in the layout of the state spectrum (main, nyq), and m1(0), m2(0) in FP64
    cpx<T>* mtab[2] = {};
    double* m0 = nullptr;
};

// dense form of the update polynomials: the smallest of 2 and 4 that covers every LAP
// power (0 when some power exceeds kDenseDeg); fills dd / dl
template <typename T>
inline int make_dense(PencilUpdateArgs<T>& u) {
    int maxp = u.src_lap;
    for (int t = 0; t < u.denom.n; ++t) maxp = maxp > u.denom.pow[t] ? maxp : u.denom.pow[t];
    for (int t = 0; t < u.lin.n; ++t) maxp = maxp > u.lin.pow[t] ? maxp : u.lin.pow[t];
    if (maxp > kDenseDeg) return u.dense = 0;
    for (int p = 0; p <= kDenseDeg; ++p) u.dd[p] = u.dl[p] = 0.0;
    for (int t = 0; t < u.denom.n; ++t) u.dd[u.denom.pow[t]] += u.denom.coef[t];
    for (int t = 0; t < u.lin.n; ++t) u.dl[u.lin.pow[t]] += u.lin.coef[t];
    return u.dense = (maxp <= 2) ? 2 : 4;
}

// x-row R2C of pairs of real arrays into main/nyq half spectra
template <typename T>
struct RowR2CArgs {
    const cpx<T>* tw = nullptr;
    const T* in[kMaxArr] = {};           // real rows (may be null -> zero)
    cpx<T>* main[kMaxArr] = {};
    cpx<T>* nyq[kMaxArr] = {};
    int narr = 0;                        // arrays processed in pairs
    long rows = 0;
    int nx = 0;
    int TW = 16;
    T in_scale = T(1);
    // optional per array: the real row is a dense polynomial of one or two fields evaluated
    // while loading (a derived field, sum_ij dc[i][j] f0^i f1^j in T, Horner in f0 of
    // polynomials in f1), so the derived field never goes through HBM
    int poly[kMaxArr] = {};
    const T* pf0[kMaxArr] = {};
    const T* pf1[kMaxArr] = {};
    double dc[kMaxArr][4][4] = {};
};

// x-row pass between stages of the general stepper: the inverse transform of one or two
// fields' rows (one transform, the fields as A + i B) gives the real rows, the dense
// derived-field polynomials (RowR2CArgs::dc, f0 = field 0, f1 = field 1) are evaluated on
// them and transformed forward in pairs into the derived fields' half spectra, so the real
// fields never go through HBM (written only where fout is set)
template <typename T>
struct RowSyncArgs {
    const cpx<T>* tw = nullptr;
    const cpx<T>* fmain[2] = {};         // fields' half spectra (fmain[1] null: one field)
    const cpx<T>* fnyq[2] = {};
    T* fout[2] = {};                     // optional real fields
    T scale = T(1);                      // inverse scale of the fields (1 / N)
    // mode 1, the elastic source of HomogeneousModulusLinElast (homoLinElast.go:91-127): one
    // array, S = the inverse row of fmain[0], and dmain[0] receives the forward row of
    // H'(phi) (S - two_estar H(phi)), phi read from the real rows of phi
    // mode 2, the adjoint's lam = p + g'(c) q (p, q the inverse rows of fmain[0], fmain[1];
    // c the real rows of phi, or g' = 0 when phi is null; g' = sum_k gd[k] c^k in FP64):
    // fout[0] receives lam and dmain[0] its forward row
    int mode = 0;
    const T* phi = nullptr;
    double two_estar = 0.0;
    int npoly = 0;
    double gd[kMaxPoly] = {};
    int nd = 0;                          // derived arrays, 1 .. kMaxArr
    double dc[kMaxArr][4][4] = {};
    cpx<T>* dmain[kMaxArr] = {};
    cpx<T>* dnyq[kMaxArr] = {};
    long rows = 0;
    int nx = 0;
    int TW = 16;
};

// x-row C2R of half spectra into real rows (scaled)
template <typename T>
struct RowC2RArgs {
    const cpx<T>* tw = nullptr;
    const cpx<T>* main[kMaxArr] = {};
    const cpx<T>* nyq[kMaxArr] = {};
    T* out[kMaxArr] = {};
    int narr = 0;
    long rows = 0;
    int nx = 0;
    int TW = 16;
    T scale = T(1);
    const double* add_mean = nullptr;    // optional FP64 value added to every output (first array)
};

// fused x-row pass of the single-field fast path:
//   x half spectra of c --C2R--> c --poly--> g --R2C--> x half spectra of g
// Each transform line carries a pair of rows (r, r + 1) as one complex sequence, in
// both directions.  The output may alias the input (each row is read whole first).
template <typename T>
struct RowFusedArgs {
    const cpx<T>* tw = nullptr;
    const cpx<T>* s_main = nullptr;
    const cpx<T>* s_nyq = nullptr;
    cpx<T>* g_main = nullptr;
    cpx<T>* g_nyq = nullptr;
    T* c_out = nullptr;                  // optional real-space copy of c
    const double* mean = nullptr;        // FP64 mean held outside the arrays (state = c - mean), or null
    long rows = 0;
    int nx = 0;
    int TW = 16;                         // transform lines (row pairs) per block
    int variant = 0;                     // transform schedule: 0 plain, 1 low-register
    int npoly = 0;                       // g(c) = sum_q gpoly[q] c^q
    double gpoly[kMaxPoly];
    int nbulk = 0;                       // optional energy density f(c) = sum_q bulk[q] c^q
    double bulk[kMaxPoly];
    double* bulk_partial = nullptr;      // per-block FP64 sum of f(c)
    double* mass_partial = nullptr;      // per-block FP64 sum of c
    // optional (monitor, blocks of >= 64 threads): block 0 also writes the previous step's
    // energy-series slot from that step's partial sums (series_slot), so that a sequence of
    // steps needs no launch of its own per slot; the previous partials are other buffers
    const double* prev_bulk = nullptr;
    const double* prev_mass = nullptr;
    const double* prev_eg = nullptr;
    long prev_nb = 0, prev_ne = 0;
    double prev_ceg = 0.0;
    double* series = nullptr;
    long* counter = nullptr;
    // batched sims: rows_per_sim > 0 (even) selects gpoly_sim[sim][kMaxPoly] and mean[sim], sim = row / rows_per_sim
    long rows_per_sim = 0;
    const double* gpoly_sim = nullptr;
};

// One field's RK4 stage (rk4.go:35-127) done as the field's values are loaded by its first
// inverse pencil pass: the stage's right-hand side from at most kStageTerms bricks of tmode 0,
// the stage factor from its table, FINAL / INIT updated in place, and the corrected spectrum
// transformed (and stored to kn when store_kn) instead of being written and read back.
constexpr int kStageTerms = 4;
template <typename T>
struct StageLine {
    const cpx<T>* k = nullptr;           // stage 0: the field spectrum (copied to init)
    cpx<T>* init = nullptr;
    cpx<T>* final = nullptr;
    cpx<T>* kn = nullptr;                // the corrected spectrum
    const T* stab = nullptr;             // the stage factor 1 / (1 - gamma dt D(k))
    int nterm = 0;
    const cpx<T>* brick[kStageTerms] = {};  // null: a constant term
    double coef[kStageTerms] = {};
    int pw[kStageTerms] = {};
    const double* dmul[kStageTerms] = {};
};
template <typename T>
struct PencilStageArgs {
    PencilArgs<T> p;                     // the inverse pencil pass; arr[q] = its outputs, narr fields
    StageLine<T> f[kMaxArr];
    int stage = 0;
    double dt = 0.0, fp = 0.0, fc = 0.0;
    int need_l = 0;                      // some term has a LAP power
    int store_kn = 1;
    KGrid g;
    long nmain = 0, nk = 0;
};
bool stage_pencil_length(int n);  // lengths the fused stage pencils are compiled for
template <typename T> hipError_t launch_pencils_stage(const PencilStageArgs<T>& a, int n, hipStream_t s);

// A run of fused 2D steps in one cooperative launch (k_fused_persist): the row pass of a
// launch's first step, of odd steps and of even steps after the first (partial-sum sets and
// series slots alternate), the update pass of even and odd steps, and the passes' block counts
// (the row pass, schedule 1, and the update pass with the same threads per block).
template <typename T>
struct PersistPlan {
    RowFusedArgs<T> ra[3];
    PencilUpdateArgs<T> u[2];
    int nblk_rows = 0, nblk_upd = 0;
    int rows_pass = 0;
    // with more participants than row blocks, the last participant (idle in the row pass)
    // writes the previous step's series slot from ra[].prev_* (whose series is then null)
    double* slot_series = nullptr;
    long* slot_counter = nullptr;
};
template <typename T>
struct PersistLaunch {
    PersistPlan<T> plan;                   // passed by value (2.4 KB of the 4 KB of kernel arguments)
    int n = 0;                             // the square grid's side: 32, 64, 128 or 256
    int mon = 0;                           // the energy monitor's partial sums
    int steps = 0;
    int local = 0;                         // participants on one XCD: the grid is 8 blocks per participant
    int blocks = 0, threads = 0;           // participating blocks
    size_t smem = 0;
    unsigned* bar = nullptr;               // 4 words, zero at the launch's start
    unsigned spin_limit = 0;               // polls of a barrier's generation before it is declared failed
    int* host_fail = nullptr;              // mapped host flag: 1 a failed barrier, 2 (local) participants
                                           // on more than one XCD, the steps not run
};
bool persist_side(int n);  // the square sides with a persistent step
// fills u.mtab and u.m0 with the update pass's multipliers (u.p set as for the pass, n = u.p.n)
template <typename T> hipError_t launch_update_mult_table(const PencilUpdateArgs<T>& u, hipStream_t s);
template <typename T> void persist_geometry(PersistLaunch<T>& L, int tw);  // threads and smem of L.n, L.mon
template <typename T> hipError_t launch_fused_persist(const PersistLaunch<T>& L, hipStream_t s);

// launchers (implemented in spectral.hip); n is the line length.
// Grouped launches of a batch of size classes: class c owns blocks [first[c], first[c] +
// blocks[c]) of the launch (first[c] a multiple of 8, first[ncls] the launch's block count);
// rows / upd point to device arrays of the classes' arguments.
constexpr int kMaxGroup = 8;
template <typename T>
struct BatchGroupArgs {
    int ncls = 0;
    int first[kMaxGroup + 1] = {};
    int blocks[kMaxGroup] = {};
    const RowFusedArgs<T>* rows = nullptr;
    const PencilUpdateArgs<T>* upd = nullptr;
};
template <typename T> hipError_t launch_rows_fused_group(const BatchGroupArgs<T>& g, int threads, size_t smem,
                                                         hipStream_t s);
template <typename T> hipError_t launch_pencil_update_group(const BatchGroupArgs<T>& g, int threads, size_t smem,
                                                            hipStream_t s);
bool group_length(int n);  // transform lengths the grouped launches are compiled for

template <typename T> hipError_t launch_pencils(const PencilArgs<T>& a, int n, int dir, hipStream_t s);
template <typename T, int DIR> hipError_t launch_pencils_dir(const PencilArgs<T>& a, int n, hipStream_t s);
template <typename T> hipError_t launch_bluestein(const BluesteinArgs<T>& a, int n, hipStream_t s);
template <typename T> hipError_t launch_pencil_update(const PencilUpdateArgs<T>& a, int n, hipStream_t s);
template <typename T> hipError_t launch_rows_r2c(const RowR2CArgs<T>& a, hipStream_t s);
template <typename T> hipError_t launch_rows_c2r(const RowC2RArgs<T>& a, hipStream_t s);
template <typename T> hipError_t launch_rows_sync(const RowSyncArgs<T>& a, hipStream_t s);
template <typename T> hipError_t launch_rows_fused(const RowFusedArgs<T>& a, hipStream_t s);

bool fft_size_supported(int n);      // sizes with a device transform
bool fft_size_fast(int n);           // sizes with a compiled Stockham schedule
int fft_threads_per_line(int n, int variant = 0);  // threads per line (variant 1: the update pass)

// blocks a launch uses (for sizing per-block partial-sum buffers)
long rows_blocks(long rows, int TW);
long rows_fused_blocks(long rows, int TW);  // the fused row pass: TW row pairs per block
long pencil_blocks(const Seg* seg, int nseg, int TW, int* tiles0);

}  // namespace gpf
