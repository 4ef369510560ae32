// CPU baseline: the reference's own CPU solvers, restated.
//
// TEST INFRASTRUCTURE.  Only tests/ and the cpu_baseline leg of bench.py load
// this file's library; nothing in the HIP product path uses it.
//
// The reference's CPU half (Tridiagonal_AVX512) solves one system per OpenMP
// thread (`omp parallel for schedule(static)` over the batch, threading.h:11-14)
// and uses the SIMD lanes of one 64-byte vector inside a system: L = 16 lanes in
// FP32, 8 in FP64.  Its solver table (Tridiagonal_AVX512/helpers.h:34-55) has
// Thomas (id 1), PCR-pThomas (id 3) and WM-pGE (id 5); these three are restated
// here from the algorithms, written with GCC vector extensions instead of
// intrinsics:
//
//   Thomas       tridiagonal_basic.cpp:6-26: scalar recurrence per system.
//   WM-pGE       tridiagonal_d_avx512.cpp:74-241 (WM3_double) and
//                tridiagonal_s_avx512.cpp:86-288 (WM4_single).  Rows are taken in
//                chunks of 2L.  log2(L) Wang-Mou merge rounds turn each chunk into
//                two blocks of L rows in which every row is coupled only to the
//                rows just outside its block.  Partial Gaussian elimination against
//                the previous block's normalised last row then writes
//                x_j = d'_j - c'_j * x_(first row of the next block) into c, d; a
//                backward sweep over the blocks finishes x.
//   PCR-pThomas  tridiagonal_d_avx512.cpp:419-534 (PCR3_double) and
//                tridiagonal_s_avx512.cpp:502-650 (PCR4_single).  A sliding window
//                of L-row blocks runs log2(L) PCR steps (strides 1, 2, ..., L/2),
//                after which lane j of block k is row k of an independent stride-L
//                subsystem; each lane runs Thomas across the blocks (c', d' into
//                c, d) and a backward sweep.
//
// Differences from the reference, all deliberate:
//   * Division is IEEE division (the reference's FP32 default is the KNL-only
//     28-bit reciprocal, vec.h:15-17,46).
//   * a_0 and c_(n-1) are treated as zero, as the FP64 oracle ignores them; the
//     reference relies on its generator storing zeros there.
//   * The PCR window keeps blocks aligned to L rows (one block of delay per PCR
//     step); the reference shifts each step's output by a fraction of a block.
// Memory traffic is the reference's: read a, b, c, d and write c', d' on the way
// down, read c', d' and write x on the way up (9 words per unknown).
//
// Requirements (the reference's, unchecked there): n % 2L == 0 for WM-pGE,
// n % L == 0 for PCR-pThomas.  Arrays of a system start 64-byte aligned in the
// reference; here any alignment works (unaligned vector loads).
//
// Built twice by oracle/__init__.py: with -mavx512f (used when the host CPU has
// AVX-512F) and for generic x86-64 (the compiler splits each 64-byte vector into
// narrower ones).
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

namespace {

template <typename T> struct Simd;
template <> struct Simd<float> {
    typedef float V __attribute__((vector_size(64)));
    typedef int32_t I __attribute__((vector_size(64)));
    static constexpr int L = 16;
    static constexpr int R = 4;   // log2(L)
};
template <> struct Simd<double> {
    typedef double V __attribute__((vector_size(64)));
    typedef int64_t I __attribute__((vector_size(64)));
    static constexpr int L = 8;
    static constexpr int R = 3;
};

template <typename V> static inline V load(const void *p) {
    V v;
    __builtin_memcpy(&v, p, sizeof v);
    return v;
}
template <typename V> static inline void store(void *p, V v) { __builtin_memcpy(p, &v, sizeof v); }

// Lane-index vectors of one Wang-Mou round (block size s = 2^(r+1), half h = s/2):
// lane j at offset t = j % s of its block reads the low half's last row
// (ll = j - t + h - 1) and the high half's first row (hf = j - t + h).
template <typename T> struct WMRound {
    typename Simd<T>::I ll, hf, high, hfirst;
};
template <typename T> struct WMConsts {
    WMRound<T> round[Simd<T>::R];
    typename Simd<T>::I zero, lane0, last;
};

template <typename T> static WMConsts<T> wm_consts() {
    constexpr int L = Simd<T>::L;
    WMConsts<T> k;
    for (int r = 0; r < Simd<T>::R; ++r) {
        const int s = 2 << r, h = s / 2;
        for (int j = 0; j < L; ++j) {
            const int t = j % s;
            k.round[r].ll[j] = j - t + h - 1;
            k.round[r].hf[j] = j - t + h;
            k.round[r].high[j] = t >= h ? -1 : 0;
            k.round[r].hfirst[j] = t == h ? -1 : 0;
        }
    }
    for (int j = 0; j < L; ++j) {
        k.zero[j] = 0;
        k.lane0[j] = j == 0 ? -1 : 0;
        k.last[j] = L - 1;
    }
    return k;
}

// One Wang-Mou round on one vector of L rows.  Row j of a block of the previous
// round reads  a_j x_p + b_j x_j + c_j x_q = d_j  with p, q the rows just outside
// that block.  Two neighbouring blocks are merged: the low block's last row
// eliminates x_ll from the high block, then the high block's first row (updated)
// eliminates x_hf from every other row of both.
template <typename T>
static inline void wm_round(typename Simd<T>::V &a, typename Simd<T>::V &b, typename Simd<T>::V &c,
                            typename Simd<T>::V &d, const WMRound<T> &m) {
    typedef typename Simd<T>::V V;
    const V zero = {};
    const V all = __builtin_shuffle(a, m.ll), bll = __builtin_shuffle(b, m.ll);
    const V cll = __builtin_shuffle(c, m.ll), dll = __builtin_shuffle(d, m.ll);
    const V k1 = m.high ? a / bll : zero;
    // high rows: coupling to x_hf created by the elimination (the pivot itself
    // for the high block's first row)
    const V e = (m.hfirst ? b : zero) - k1 * cll;
    a = m.high ? -(k1 * all) : a;
    b = m.hfirst ? e : b;
    d = d - k1 * dll;
    const V ahf = __builtin_shuffle(a, m.hf), bhf = __builtin_shuffle(b, m.hf);
    const V chf = __builtin_shuffle(c, m.hf), dhf = __builtin_shuffle(d, m.hf);
    const V k2 = (m.hfirst ? zero : (m.high ? e : c)) / bhf;
    a = a - k2 * ahf;
    c = m.high ? c - k2 * chf : -(k2 * chf);
    d = d - k2 * dhf;
}

// Partial Gaussian elimination of one merged block of L rows against the
// previous block's normalised last row x_p = ld - lc * x_(this block's row 0).
// Leaves every row normalised: x_j = d_j - c_j * x_(next block's row 0).
template <typename T>
static inline void wm_pge(typename Simd<T>::V &a, typename Simd<T>::V &b, typename Simd<T>::V &c,
                          typename Simd<T>::V &d, typename Simd<T>::V lc, typename Simd<T>::V ld,
                          const WMConsts<T> &k) {
    typedef typename Simd<T>::V V;
    const V zero = {};
    const V e = (k.lane0 ? b : zero) - a * lc;
    d = d - a * ld;
    b = k.lane0 ? e : b;
    const V bf = __builtin_shuffle(e, k.zero), cf = __builtin_shuffle(c, k.zero);
    const V df = __builtin_shuffle(d, k.zero);
    const V q = (k.lane0 ? zero : e) / bf;
    c = (c - q * cf) / b;
    d = (d - q * df) / b;
}

template <typename T>
static void wm_pge_system(const T *a, const T *b, T *c, T *d, int n, const WMConsts<T> &k) {
    typedef typename Simd<T>::V V;
    constexpr int L = Simd<T>::L;
    V lc = {}, ld = {};
    for (int i = 0; i < n; i += 2 * L) {
        V a0 = load<V>(a + i), b0 = load<V>(b + i), c0 = load<V>(c + i), d0 = load<V>(d + i);
        V a1 = load<V>(a + i + L), b1 = load<V>(b + i + L), c1 = load<V>(c + i + L),
          d1 = load<V>(d + i + L);
        if (i == 0) a0[0] = 0;
        if (i + 2 * L == n) c1[L - 1] = 0;
#pragma GCC unroll 4
        for (int r = 0; r < Simd<T>::R; ++r) {
            wm_round<T>(a0, b0, c0, d0, k.round[r]);
            wm_round<T>(a1, b1, c1, d1, k.round[r]);
        }
        wm_pge<T>(a0, b0, c0, d0, lc, ld, k);
        store(c + i, c0);
        store(d + i, d0);
        lc = __builtin_shuffle(c0, k.last);
        ld = __builtin_shuffle(d0, k.last);
        wm_pge<T>(a1, b1, c1, d1, lc, ld, k);
        store(c + i + L, c1);
        store(d + i + L, d1);
        lc = __builtin_shuffle(c1, k.last);
        ld = __builtin_shuffle(d1, k.last);
    }
    // the last block couples to nothing on its right (c_(n-1) = 0): its d' is x
    V xf = __builtin_shuffle(load<V>(d + n - L), k.zero);
    for (int i = n - 2 * L; i >= 0; i -= L) {
        const V x = load<V>(d + i) - load<V>(c + i) * xf;
        store(d + i, x);
        xf = __builtin_shuffle(x, k.zero);
    }
}

// ------------------------------------------------------------------ PCR-pThomas
// Two-vector shuffle indices of one PCR step with stride s: lane j of block k
// reads row kL + j - s from (block k-1 | block k) and row kL + j + s from
// (block k | block k+1).
template <typename T> struct PCRConsts {
    typename Simd<T>::I lo[Simd<T>::R], hi[Simd<T>::R];
};
template <typename T> static PCRConsts<T> pcr_consts() {
    constexpr int L = Simd<T>::L;
    PCRConsts<T> k;
    for (int r = 0; r < Simd<T>::R; ++r) {
        const int s = 1 << r;
        for (int j = 0; j < L; ++j) {
            k.lo[r][j] = L + j - s;
            k.hi[r][j] = j + s;
        }
    }
    return k;
}

template <typename T> struct Rows {
    typename Simd<T>::V a, b, c, d;
};

template <typename T> static inline Rows<T> identity_rows() {
    Rows<T> x;
    x.a = x.c = x.d = typename Simd<T>::V{};
    x.b = x.a + (T)1;
    return x;
}

// One PCR step (stride s) on block k, given blocks k-1 and k+1 of its input.
template <typename T>
static inline Rows<T> pcr_step(const Rows<T> &prev, const Rows<T> &cur, const Rows<T> &next,
                               const typename Simd<T>::I &lo, const typename Simd<T>::I &hi) {
    typedef typename Simd<T>::V V;
    const V al = __builtin_shuffle(prev.a, cur.a, lo), bl = __builtin_shuffle(prev.b, cur.b, lo);
    const V cl = __builtin_shuffle(prev.c, cur.c, lo), dl = __builtin_shuffle(prev.d, cur.d, lo);
    const V ah = __builtin_shuffle(cur.a, next.a, hi), bh = __builtin_shuffle(cur.b, next.b, hi);
    const V ch = __builtin_shuffle(cur.c, next.c, hi), dh = __builtin_shuffle(cur.d, next.d, hi);
    const V k1 = cur.a / bl, k2 = cur.c / bh;
    Rows<T> o;
    o.a = -(k1 * al);
    o.c = -(k2 * ch);
    o.b = cur.b - k1 * cl - k2 * ah;
    o.d = cur.d - k1 * dl - k2 * dh;
    return o;
}

template <typename T>
static void pcr_pthomas_system(const T *a, const T *b, T *c, T *d, int n, const PCRConsts<T> &k) {
    typedef typename Simd<T>::V V;
    constexpr int L = Simd<T>::L, R = Simd<T>::R;
    const int nb = n / L;
    const Rows<T> id = identity_rows<T>();
    auto input = [&](int t) {
        if (t >= nb) return id;
        Rows<T> x{load<V>(a + t * L), load<V>(b + t * L), load<V>(c + t * L), load<V>(d + t * L)};
        if (t == 0) x.a[0] = 0;
        if (t == nb - 1) x.c[L - 1] = 0;
        return x;
    };
    // step r holds the two latest blocks of its input; it emits block k once block
    // k+1 has arrived, so step r runs r+1 blocks behind the loads
    Rows<T> prev[R], cur[R];
    for (int t = 0; t < R; ++t) {
        Rows<T> v = input(t);
        for (int r = 0; r < t; ++r) {
            const Rows<T> o = pcr_step<T>(prev[r], cur[r], v, k.lo[r], k.hi[r]);
            prev[r] = cur[r];
            cur[r] = v;
            v = o;
        }
        prev[t] = id;
        cur[t] = v;
    }
    V lc = {}, ld = {};
    for (int t = R; t < nb + R; ++t) {
        Rows<T> v = input(t);
#pragma GCC unroll 4
        for (int r = 0; r < R; ++r) {
            const Rows<T> o = pcr_step<T>(prev[r], cur[r], v, k.lo[r], k.hi[r]);
            prev[r] = cur[r];
            cur[r] = v;
            v = o;
        }
        // lane j of block t-R is row t-R of stride-L subsystem j: one Thomas step
        const V piv = v.b - lc * v.a;
        lc = v.c / piv;
        ld = (v.d - ld * v.a) / piv;
        store(c + (t - R) * L, lc);
        store(d + (t - R) * L, ld);
    }
    V x = load<V>(d + (nb - 1) * L);
    for (int t = nb - 2; t >= 0; --t) {
        x = load<V>(d + t * L) - load<V>(c + t * L) * x;
        store(d + t * L, x);
    }
}

// ------------------------------------------------------------------ Thomas
template <typename T> static void thomas_system(const T *a, const T *b, T *c, T *d, int n) {
    c[0] = c[0] / b[0];
    d[0] = d[0] / b[0];
    for (int i = 1; i < n; ++i) {
        const T k = b[i] - c[i - 1] * a[i];
        c[i] = c[i] / k;
        d[i] = (d[i] - d[i - 1] * a[i]) / k;
    }
    for (int i = n - 2; i >= 0; --i) d[i] = d[i] - c[i] * d[i + 1];
}

template <typename T>
static int solve_batch(int algo, int n, int batch, const T *a, const T *b, T *c, T *d) {
    constexpr int L = Simd<T>::L;
    if (n < 1 || batch < 0) return 1;
    if (algo == 1) {
#pragma omp parallel for schedule(static)
        for (int s = 0; s < batch; ++s)
            thomas_system<T>(a + (int64_t)s * n, b + (int64_t)s * n, c + (int64_t)s * n,
                             d + (int64_t)s * n, n);
        return 0;
    }
    if (algo == 3) {
        if (n % L) return 1;
        const PCRConsts<T> k = pcr_consts<T>();
#pragma omp parallel for schedule(static)
        for (int s = 0; s < batch; ++s)
            pcr_pthomas_system<T>(a + (int64_t)s * n, b + (int64_t)s * n, c + (int64_t)s * n,
                                  d + (int64_t)s * n, n, k);
        return 0;
    }
    if (algo == 5) {
        if (n % (2 * L)) return 1;
        const WMConsts<T> k = wm_consts<T>();
#pragma omp parallel for schedule(static)
        for (int s = 0; s < batch; ++s)
            wm_pge_system<T>(a + (int64_t)s * n, b + (int64_t)s * n, c + (int64_t)s * n,
                             d + (int64_t)s * n, n, k);
        return 0;
    }
    return 2;
}

template <typename T>
static double bench(int algo, int n, int batch, const T *a0, const T *b0, const T *c0,
                    const T *d0, T *x_out, int nsets, int rounds, double *t_min) {
    const size_t elems = (size_t)n * batch;
    const size_t bytes = ((elems * sizeof(T)) + 63) & ~(size_t)63;
    T **sets = (T **)calloc((size_t)nsets * 4, sizeof(T *));
    if (!sets) return -1.0;
    for (int k = 0; k < nsets * 4; ++k) {
        sets[k] = (T *)aligned_alloc(64, bytes);
        if (!sets[k]) {
            for (int j = 0; j < k; ++j) free(sets[j]);
            free(sets);
            return -1.0;
        }
    }
    const T *src[4] = {a0, b0, c0, d0};
    double total = 0.0, best = 1e30;
    int count = 0, rc = 0;
    for (int round = 0; round < rounds && rc == 0; ++round) {
        // untimed: restore every set, each by the thread that solves its systems
        for (int k = 0; k < nsets; ++k)
            for (int q = 0; q < 4; ++q) {
                T *dst = sets[4 * k + q];
                const T *sp = src[q];
#pragma omp parallel for schedule(static)
                for (int s = 0; s < batch; ++s)
                    memcpy(dst + (size_t)s * n, sp + (size_t)s * n, (size_t)n * sizeof(T));
            }
        // timed: each set once, in order; every set was last touched by the restore,
        // with the other sets' traffic since
        for (int k = 0; k < nsets && rc == 0; ++k) {
            T **st = sets + 4 * k;
            const double t0 = omp_get_wtime();
            rc = solve_batch<T>(algo, n, batch, st[0], st[1], st[2], st[3]);
            const double dt = omp_get_wtime() - t0;
            total += dt;
            best = dt < best ? dt : best;
            ++count;
        }
    }
    if (rc == 0 && x_out) memcpy(x_out, sets[4 * (nsets - 1) + 3], elems * sizeof(T));
    for (int k = 0; k < nsets * 4; ++k) free(sets[k]);
    free(sets);
    if (rc != 0) return -(double)rc;
    if (t_min) *t_min = best;
    return total / count;
}

}  // namespace

extern "C" {

// 1 when this build uses 512-bit vectors (AVX-512F), else 0.
int cpuref_avx512(void) {
#ifdef __AVX512F__
    return 1;
#else
    return 0;
#endif
}

// Solves a batch in the reference layout (system s at s*n), in place: x into d,
// c overwritten.  algo: 1 Thomas, 3 PCR-pThomas, 5 WM-pGE (the reference's CPU
// table ids).  threads > 0 sets the OpenMP thread count.  Returns 0, 1 for sizes
// the algorithm does not take, 2 for an unknown algorithm.
int cpuref_solve(int prec, int algo, int n, int batch, const void *a, const void *b, void *c,
                 void *d, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    if (prec == 0)
        return solve_batch<float>(algo, n, batch, (const float *)a, (const float *)b, (float *)c,
                                  (float *)d);
    if (prec == 1)
        return solve_batch<double>(algo, n, batch, (const double *)a, (const double *)b,
                                   (double *)c, (double *)d);
    return 2;
}

// Times `rounds` x `nsets` solves of copies of one input batch (a0..d0, unchanged)
// the way the reference's driver times a solver (host clock around each call,
// inputs restored untimed before it; Tridiagonal_AVX512/test.cpp:100-106), with
// `nsets` copies so that each solve starts from memory that the other sets have
// since pushed out of the host caches.  Writes x of the last solve to x_out (if
// not null) and the fastest solve to *t_min.  Returns seconds per solve, or a
// negative value on error.
double cpuref_bench(int prec, int algo, int n, int batch, const void *a0, const void *b0,
                    const void *c0, const void *d0, void *x_out, int nsets, int rounds,
                    int threads, double *t_min) {
    if (threads > 0) omp_set_num_threads(threads);
    if (nsets < 1 || rounds < 1 || n < 1 || batch < 1) return -1.0;
    if (prec == 0)
        return bench<float>(algo, n, batch, (const float *)a0, (const float *)b0,
                            (const float *)c0, (const float *)d0, (float *)x_out, nsets, rounds,
                            t_min);
    if (prec == 1)
        return bench<double>(algo, n, batch, (const double *)a0, (const double *)b0,
                             (const double *)c0, (const double *)d0, (double *)x_out, nsets,
                             rounds, t_min);
    return -2.0;
}

}  // extern "C"
