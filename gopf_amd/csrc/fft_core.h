// fft_core.h -- device-side building blocks of the spectral step (gfx950).
//
// Complex lines are transformed in LDS with a self-sorting (Stockham) mixed-
// radix schedule.  Each thread of a line owns E elements per stage; a stage of
// radix R runs E/R radix-R butterflies per thread entirely in registers, and
// stages exchange data through LDS.  Transform convention (FFTW's, which the
// reference calls through pfutil/fftWrap.go:19-20):
//   forward (DIR = -1): X[k] = sum_n x[n] exp(-2 pi i n k / N)
//   inverse (DIR = +1): x[n] = sum_k X[k] exp(+2 pi i n k / N)   (unnormalised)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpf {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> mk(T a, T b) { cpx<T> r; r.x = a; r.y = b; return r; }
template <typename T>
__device__ __forceinline__ cpx<T> add(cpx<T> a, cpx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> sub(cpx<T> a, cpx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> mul(cpx<T> a, cpx<T> b) {
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cpx<T> scale(cpx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
template <typename T>
__device__ __forceinline__ cpx<T> conjg(cpx<T> a) { return mk<T>(a.x, -a.y); }
// multiply by (DIR * i): DIR=-1 -> -i, DIR=+1 -> +i
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> mul_i(cpx<T> a) {
    return (DIR < 0) ? mk<T>(a.y, -a.x) : mk<T>(-a.y, a.x);
}

// ---------------------------------------------------------------------------
// compile-time radix schedules
// ---------------------------------------------------------------------------
// V = 0: default schedule; V = 1: low-register schedule (E <= 12) for kernels
// that keep several transformed lines in registers at once
template <int N, int V = 0>
struct Sched : Sched<N, 0> {};
#define GPF_SCHED_V(n, v, e, ...)                           \
    template <>                                             \
    struct Sched<n, v> {                                    \
        static constexpr int E = e;                         \
        static constexpr int R[] = {__VA_ARGS__};           \
        static constexpr int NR = sizeof(R) / sizeof(int);  \
        static constexpr int T = n / e;                     \
    };
#define GPF_SCHED(n, e, ...) GPF_SCHED_V(n, 0, e, __VA_ARGS__)
GPF_SCHED(2, 2, 2)
GPF_SCHED(4, 4, 4)
GPF_SCHED(8, 8, 8)
GPF_SCHED(16, 16, 16)
GPF_SCHED(32, 8, 8, 4)
GPF_SCHED(64, 8, 8, 8)
GPF_SCHED(128, 16, 8, 16)
GPF_SCHED(256, 16, 16, 16)
GPF_SCHED(512, 8, 8, 8, 8)
GPF_SCHED(1024, 16, 16, 16, 4)
GPF_SCHED(12, 12, 4, 3)
GPF_SCHED(24, 24, 8, 3)
GPF_SCHED(48, 12, 4, 4, 3)
GPF_SCHED(96, 24, 8, 4, 3)
GPF_SCHED(192, 24, 8, 8, 3)
GPF_SCHED(384, 24, 8, 8, 2, 3)
GPF_SCHED(768, 48, 16, 16, 3)
GPF_SCHED_V(128, 1, 8, 8, 4, 4)
GPF_SCHED_V(256, 1, 8, 8, 8, 4)
GPF_SCHED_V(1024, 1, 8, 8, 8, 4, 4)
GPF_SCHED_V(24, 1, 12, 4, 2, 3)
GPF_SCHED_V(96, 1, 12, 4, 4, 2, 3)
GPF_SCHED_V(192, 1, 12, 4, 4, 4, 3)
GPF_SCHED_V(384, 1, 12, 4, 4, 4, 2, 3)
GPF_SCHED_V(768, 1, 12, 4, 4, 4, 4, 3)
// radix-5 lengths (E = 20 or 10: a multiple of every radix of the schedule)
GPF_SCHED(80, 20, 4, 4, 5)
GPF_SCHED(100, 20, 4, 5, 5)
GPF_SCHED(160, 20, 4, 4, 2, 5)
GPF_SCHED(200, 20, 4, 2, 5, 5)
GPF_SCHED(250, 10, 2, 5, 5, 5)
GPF_SCHED(320, 20, 4, 4, 4, 5)
GPF_SCHED(400, 20, 4, 4, 5, 5)
GPF_SCHED(500, 20, 4, 5, 5, 5)
GPF_SCHED(640, 20, 4, 4, 4, 2, 5)
GPF_SCHED(800, 20, 4, 4, 2, 5, 5)
GPF_SCHED(1000, 20, 4, 2, 5, 5, 5)
GPF_SCHED(1280, 20, 4, 4, 4, 4, 5)
GPF_SCHED(1600, 20, 4, 4, 4, 5, 5)
GPF_SCHED(2000, 20, 4, 4, 5, 5, 5)
GPF_SCHED(2560, 20, 4, 4, 4, 4, 2, 5)
// radix-7 lengths
GPF_SCHED(112, 28, 4, 4, 7)
GPF_SCHED(224, 28, 4, 4, 2, 7)
GPF_SCHED(448, 28, 4, 4, 4, 7)
GPF_SCHED(896, 28, 4, 4, 4, 2, 7)
// 2^11 (the largest power of two with a compiled schedule)
GPF_SCHED(2048, 16, 16, 16, 8)
GPF_SCHED_V(2048, 1, 8, 8, 8, 8, 4)
#undef GPF_SCHED
#undef GPF_SCHED_V

// every length with a compiled schedule, for dispatch tables: X(n) per length.  The core
// lengths (2^a <= 1024, 2^a 3 <= 768) and the extended ones (2048, radix 5 and 7) are
// compiled into separate units, so that the code objects of the hot passes stay small
#define GPF_CORE_LENGTHS(X)                                                                        \
    X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)                                  \
    X(12) X(24) X(48) X(96) X(192) X(384) X(768)
#define GPF_EXT_LENGTHS(X)                                                                         \
    X(2048) X(80) X(100) X(160) X(200) X(250) X(320) X(400) X(500) X(640) X(800) X(1000) X(1280)   \
    X(1600) X(2000) X(2560) X(112) X(224) X(448) X(896)
#define GPF_FAST_LENGTHS(X) GPF_CORE_LENGTHS(X) GPF_EXT_LENGTHS(X)

// span before stage S
template <int N, int V, int S>
struct Span {
    static constexpr int v = Span<N, V, S - 1>::v * Sched<N, V>::R[S - 1];
};
template <int N, int V>
struct Span<N, V, 0> {
    static constexpr int v = 1;
};

// ---------------------------------------------------------------------------
// small DFTs in registers
// ---------------------------------------------------------------------------
template <int DIR, typename T>
__device__ __forceinline__ void dft2(cpx<T>& a0, cpx<T>& a1) {
    cpx<T> t = sub(a0, a1);
    a0 = add(a0, a1);
    a1 = t;
}

template <int DIR, typename T>
__device__ __forceinline__ void dft4(cpx<T>& a0, cpx<T>& a1, cpx<T>& a2, cpx<T>& a3) {
    cpx<T> t0 = add(a0, a2), t1 = sub(a0, a2);
    cpx<T> t2 = add(a1, a3), t3 = mul_i<DIR>(sub(a1, a3));
    a0 = add(t0, t2);
    a2 = sub(t0, t2);
    a1 = add(t1, t3);
    a3 = sub(t1, t3);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft3(cpx<T>* a) {
    const T c = T(-0.5);
    const T s = T(DIR) * T(0.86602540378443864676372317075294);  // DIR * sqrt(3)/2
    cpx<T> sm = add(a[1], a[2]), df = sub(a[1], a[2]);
    cpx<T> b = mk<T>(a[0].x + c * sm.x, a[0].y + c * sm.y);
    // i * s * df
    cpx<T> r = mk<T>(-s * df.y, s * df.x);
    a[0] = add(a[0], sm);
    a[1] = add(b, r);
    a[2] = sub(b, r);
}

// length 5: X_k = x_0 + sum_{j=1,2} (x_j + x_{5-j}) cos(2 pi jk/5)
//               + DIR i sum_{j=1,2} (x_j - x_{5-j}) sin(2 pi jk/5)
template <int DIR, typename T>
__device__ __forceinline__ void dft5(cpx<T>* a) {
    const T c1 = T(0.30901699437494742410229341718281906);   // cos(2 pi / 5)
    const T c2 = T(-0.80901699437494742410229341718281906);  // cos(4 pi / 5)
    const T s1 = T(DIR) * T(0.95105651629515357211643933337938214);  // DIR sin(2 pi / 5)
    const T s2 = T(DIR) * T(0.58778525229247312916870595463907277);  // DIR sin(4 pi / 5)
    const cpx<T> t1 = add(a[1], a[4]), t2 = add(a[2], a[3]);
    const cpx<T> u1 = sub(a[1], a[4]), u2 = sub(a[2], a[3]);
    const cpx<T> x0 = a[0];
    a[0] = add(x0, add(t1, t2));
    const cpx<T> p1 = mk<T>(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
    const cpx<T> p2 = mk<T>(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
    // i (s1 u1 + s2 u2) and i (s2 u1 - s1 u2)
    const cpx<T> q1 = mk<T>(-(s1 * u1.y + s2 * u2.y), s1 * u1.x + s2 * u2.x);
    const cpx<T> q2 = mk<T>(-(s2 * u1.y - s1 * u2.y), s2 * u1.x - s1 * u2.x);
    a[1] = add(p1, q1);
    a[4] = sub(p1, q1);
    a[2] = add(p2, q2);
    a[3] = sub(p2, q2);
}

// length 7, the same pairing of x_j and x_{7-j}; cos(2 pi m/7) and sin(2 pi m/7) for
// m = jk mod 7 reduce to c1..c3 and +-s1..s3
template <int DIR, typename T>
__device__ __forceinline__ void dft7(cpx<T>* a) {
    const T c1 = T(0.62348980185873353052500488400423981);   // cos(2 pi / 7)
    const T c2 = T(-0.22252093395631440428890256449679476);  // cos(4 pi / 7)
    const T c3 = T(-0.90096886790241912623610231950744505);  // cos(6 pi / 7)
    const T s1 = T(DIR) * T(0.78183148246802980870844452667405775);  // DIR sin(2 pi / 7)
    const T s2 = T(DIR) * T(0.97492791218182360701813168299393122);  // DIR sin(4 pi / 7)
    const T s3 = T(DIR) * T(0.43388373911755812047576833284835875);  // DIR sin(6 pi / 7)
    const cpx<T> t1 = add(a[1], a[6]), t2 = add(a[2], a[5]), t3 = add(a[3], a[4]);
    const cpx<T> u1 = sub(a[1], a[6]), u2 = sub(a[2], a[5]), u3 = sub(a[3], a[4]);
    const cpx<T> x0 = a[0];
    a[0] = add(x0, add(t1, add(t2, t3)));
    // k = 1: (c1, c2, c3), (s1, s2, s3); k = 2: (c2, c3, c1), (s2, -s3, -s1); k = 3: (c3, c1, c2), (s3, -s1, s2)
    const cpx<T> p1 = mk<T>(x0.x + c1 * t1.x + c2 * t2.x + c3 * t3.x, x0.y + c1 * t1.y + c2 * t2.y + c3 * t3.y);
    const cpx<T> p2 = mk<T>(x0.x + c2 * t1.x + c3 * t2.x + c1 * t3.x, x0.y + c2 * t1.y + c3 * t2.y + c1 * t3.y);
    const cpx<T> p3 = mk<T>(x0.x + c3 * t1.x + c1 * t2.x + c2 * t3.x, x0.y + c3 * t1.y + c1 * t2.y + c2 * t3.y);
    const T r1x = s1 * u1.x + s2 * u2.x + s3 * u3.x, r1y = s1 * u1.y + s2 * u2.y + s3 * u3.y;
    const T r2x = s2 * u1.x - s3 * u2.x - s1 * u3.x, r2y = s2 * u1.y - s3 * u2.y - s1 * u3.y;
    const T r3x = s3 * u1.x - s1 * u2.x + s2 * u3.x, r3y = s3 * u1.y - s1 * u2.y + s2 * u3.y;
    a[1] = mk<T>(p1.x - r1y, p1.y + r1x);
    a[6] = mk<T>(p1.x + r1y, p1.y - r1x);
    a[2] = mk<T>(p2.x - r2y, p2.y + r2x);
    a[5] = mk<T>(p2.x + r2y, p2.y - r2x);
    a[3] = mk<T>(p3.x - r3y, p3.y + r3x);
    a[4] = mk<T>(p3.x + r3y, p3.y - r3x);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft8(cpx<T>* a) {
    const T h = T(0.70710678118654752440084436210485);
    cpx<T> e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    cpx<T> o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    // twiddles W8^k, W8 = exp(DIR * 2 pi i / 8)
    cpx<T> w1o = mk<T>(h * (o1.x - T(DIR) * o1.y), h * (o1.y + T(DIR) * o1.x));
    cpx<T> w2o = mul_i<DIR>(o2);
    cpx<T> w3o = mk<T>(h * (-o3.x - T(DIR) * o3.y), h * (-o3.y + T(DIR) * o3.x));
    a[0] = add(e0, o0);
    a[4] = sub(e0, o0);
    a[1] = add(e1, w1o);
    a[5] = sub(e1, w1o);
    a[2] = add(e2, w2o);
    a[6] = sub(e2, w2o);
    a[3] = add(e3, w3o);
    a[7] = sub(e3, w3o);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft16(cpx<T>* a) {
    // 4 x 4 split: m = m2 + 4 m1, k = k1 + 4 k2
    cpx<T> b[16];
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) {
        cpx<T> x0 = a[m2], x1 = a[m2 + 4], x2 = a[m2 + 8], x3 = a[m2 + 12];
        dft4<DIR>(x0, x1, x2, x3);
        b[m2 * 4 + 0] = x0;
        b[m2 * 4 + 1] = x1;
        b[m2 * 4 + 2] = x2;
        b[m2 * 4 + 3] = x3;
    }
    // twiddle W16^(m2 k1)
    const T c1 = T(0.92387953251128675612818318939679);
    const T s1 = T(0.38268343236508977172845998403040);
    const T h = T(0.70710678118654752440084436210485);
    const T sg = T(DIR);
    // W16^e for e in 1..9 with W16 = exp(DIR 2 pi i / 16)
    const cpx<T> w1 = mk<T>(c1, sg * s1), w2 = mk<T>(h, sg * h), w3 = mk<T>(s1, sg * c1);
    const cpx<T> w6 = mk<T>(-h, sg * h), w9 = mk<T>(-c1, -sg * s1);
    b[1 * 4 + 1] = mul(b[1 * 4 + 1], w1);
    b[1 * 4 + 2] = mul(b[1 * 4 + 2], w2);
    b[1 * 4 + 3] = mul(b[1 * 4 + 3], w3);
    b[2 * 4 + 1] = mul(b[2 * 4 + 1], w2);
    b[2 * 4 + 2] = mul_i<DIR>(b[2 * 4 + 2]);
    b[2 * 4 + 3] = mul(b[2 * 4 + 3], w6);
    b[3 * 4 + 1] = mul(b[3 * 4 + 1], w3);
    b[3 * 4 + 2] = mul(b[3 * 4 + 2], w6);
    b[3 * 4 + 3] = mul(b[3 * 4 + 3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cpx<T> x0 = b[0 * 4 + k1], x1 = b[1 * 4 + k1], x2 = b[2 * 4 + k1], x3 = b[3 * 4 + k1];
        dft4<DIR>(x0, x1, x2, x3);
        a[k1] = x0;
        a[k1 + 4] = x1;
        a[k1 + 8] = x2;
        a[k1 + 12] = x3;
    }
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dftR(cpx<T>* a) {
    if constexpr (R == 2) {
        dft2<DIR>(a[0], a[1]);
    } else if constexpr (R == 3) {
        dft3<DIR>(a);
    } else if constexpr (R == 4) {
        dft4<DIR>(a[0], a[1], a[2], a[3]);
    } else if constexpr (R == 5) {
        dft5<DIR>(a);
    } else if constexpr (R == 7) {
        dft7<DIR>(a);
    } else if constexpr (R == 8) {
        dft8<DIR>(a);
    } else if constexpr (R == 16) {
        dft16<DIR>(a);
    }
}

// ---------------------------------------------------------------------------
// register-resident transform of one line
//
// "Canonical" register order: thread t (of T per line) holds v[q] = x[t + T q],
// q in [0, E).  Stage s of radix R runs E/R butterflies per thread; butterfly
// b has index j = t + b T, reads x[j + m N/R] and writes
// y[(j / NS) NS R + (j % NS) + m NS] (Stockham, NS = span before the stage).
// The first stage reads canonical registers directly and the last stage
// produces canonical registers, so loads and stores of whole lines need no
// LDS staging; stages exchange through LDS (padded one slot per 16).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lds_pad(int p) { return p + (p >> 4); }
constexpr int lds_line(int n) { return n + n / 16 + 1; }

// twiddles of stage S: butterfly b, leg m (1 <= m < R) takes w^(k m TWSTEP), k = j mod NS
template <int N, int V, int S>
struct StageTw {
    static constexpr int R = Sched<N, V>::R[S];
    static constexpr int NB = Sched<N, V>::E / R;
    static constexpr int NS = Span<N, V, S>::v;
    static constexpr int TWSTEP = N / (NS * R);
    static constexpr int count = (NS > 1) ? NB * (R - 1) : 1;
};

template <typename T, int N, int V, int S>
__device__ __forceinline__ void load_stage_tw(cpx<T>* w, int t, const cpx<T>* __restrict__ tw) {
    using ST = StageTw<N, V, S>;
    if constexpr (ST::NS > 1) {
        constexpr int TT = Sched<N, V>::T;
#pragma unroll
        for (int b = 0; b < ST::NB; ++b) {
            const int k = (t + b * TT) % ST::NS;
#pragma unroll
            for (int m = 1; m < ST::R; ++m) w[b * (ST::R - 1) + m - 1] = tw[(k * m * ST::TWSTEP) % N];
        }
    }
}

// stage S's butterflies on registers a, with the stage's twiddles w (load_stage_tw)
template <typename T, int N, int V, int S, int DIR>
__device__ __forceinline__ void stage_compute(cpx<T>* a, const cpx<T>* w) {
    using ST = StageTw<N, V, S>;
#pragma unroll
    for (int b = 0; b < ST::NB; ++b) {
        if constexpr (ST::NS > 1) {
#pragma unroll
            for (int m = 1; m < ST::R; ++m) {
                cpx<T> x = w[b * (ST::R - 1) + m - 1];
                if (DIR > 0) x.y = -x.y;
                a[b * ST::R + m] = mul(a[b * ST::R + m], x);
            }
        }
        dftR<ST::R, DIR>(&a[b * ST::R]);
    }
}

template <typename T, int N, int V, int S, int DIR>
__device__ __forceinline__ void exchange_and_compute(cpx<T>* a, cpx<T>* lds, int es, int t,
                                                     const cpx<T>* __restrict__ tw) {
    if constexpr (S < Sched<N, V>::NR) {
        constexpr int TT = Sched<N, V>::T;
        constexpr int E = Sched<N, V>::E;
        constexpr int RP = Sched<N, V>::R[S - 1];
        constexpr int NBP = E / RP;
        constexpr int NSP = Span<N, V, S - 1>::v;
        constexpr int R = Sched<N, V>::R[S];
        constexpr int NB = E / R;
        // the stage's twiddles are requested before the exchange, so that their latency
        // overlaps the LDS round trip and the barriers
        cpx<T> w[StageTw<N, V, S>::count];
        load_stage_tw<T, N, V, S>(w, t, tw);
#pragma unroll
        for (int b = 0; b < NBP; ++b) {
            const int j = t + b * TT;
            const int base = (NSP == 1) ? j * RP : ((j / NSP) * NSP * RP + (j % NSP));
#pragma unroll
            for (int m = 0; m < RP; ++m) lds[lds_pad(base + m * NSP) * es] = a[b * RP + m];
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * TT;
#pragma unroll
            for (int m = 0; m < R; ++m) a[b * R + m] = lds[lds_pad(j + m * (N / R)) * es];
        }
        __syncthreads();
        stage_compute<T, N, V, S, DIR>(a, w);
        exchange_and_compute<T, N, V, S + 1, DIR>(a, lds, es, t, tw);
    }
}

// In-place transform of a line held in canonical registers v[E].  lds is this
// line's exchange buffer (element p at lds[lds_pad(p) * es]).  Every thread of
// the block must call it: it synchronises the block between stages.
template <typename T, int N, int DIR, int V = 0>
__device__ __forceinline__ void fft_canon(cpx<T>* v, cpx<T>* lds, int es, int t, const cpx<T>* __restrict__ tw) {
    constexpr int E = Sched<N, V>::E;
    constexpr int NR = Sched<N, V>::NR;
    constexpr int R0 = Sched<N, V>::R[0];
    constexpr int NB0 = E / R0;
    constexpr int RL = Sched<N, V>::R[NR - 1];
    constexpr int NBL = E / RL;
    cpx<T> a[E];
#pragma unroll
    for (int b = 0; b < NB0; ++b)
#pragma unroll
        for (int m = 0; m < R0; ++m) a[b * R0 + m] = v[b + m * NB0];
    {
        cpx<T> w[StageTw<N, V, 0>::count];
        load_stage_tw<T, N, V, 0>(w, t, tw);
        stage_compute<T, N, V, 0, DIR>(a, w);
    }
    exchange_and_compute<T, N, V, 1, DIR>(a, lds, es, t, tw);
#pragma unroll
    for (int b = 0; b < NBL; ++b)
#pragma unroll
        for (int m = 0; m < RL; ++m) v[b + m * NBL] = a[b * RL + m];
}

// Brute-force DFT for runtime lengths n <= 64: one thread per element (T = n, E = 1).
template <typename T, int DIR>
__device__ __forceinline__ void dft_canon_rt(cpx<T>* v, cpx<T>* lds, int es, int n, int t,
                                             const cpx<T>* __restrict__ tw) {
    lds[lds_pad(t) * es] = v[0];
    __syncthreads();
    cpx<T> acc = mk<T>(T(0), T(0));
    int q = 0;
    for (int m = 0; m < n; ++m) {
        cpx<T> w = tw[q];
        if (DIR > 0) w.y = -w.y;
        acc = add(acc, mul(lds[lds_pad(m) * es], w));
        q += t;
        if (q >= n) q -= n;
    }
    __syncthreads();
    v[0] = acc;
}

}  // namespace gpf
