// Bulk-row forms of the reference's streaming solvers: one wave streams one
// system through its 64 lanes in row blocks, with the reference's memory
// semantics (c' into c and d' into d on the way down, x into d on the way up:
// 9 words per unknown).  These are the algorithms of the reference's ids 3-10
// themselves, for callers that want them rather than the partition kernel that
// tridiag_solve runs for those ids by default.
//
//   WM-pGE (Tridiagonal_CUDA/tridiagonal_wm_pcr.cu:17-205, 424-646; WM-half-pGE
//   :207-422, 648-884).  A chunk is 128 rows: lane t holds rows t and 64 + t.
//   Every row starts in outer form relative to its own neighbours
//   (x_i = d_i/b_i - a_i/b_i x_{i-1} - c_i/b_i x_{i+1}); ROUNDS Wang-Mou merge
//   rounds over the lanes double the blocks up to 2^ROUNDS rows (64 for the
//   full form, 32 for the half form), after which every row is coupled only to
//   the rows just outside its block.  Partial Gaussian elimination against the
//   previous block's last row, x_E = rho - sigma x_F, then gives every row as
//   x_i = d'_i - c'_i x_(row after the block), stored into c and d.  The way
//   back runs block by block from the end.
//
//   PCR-pThomas (tridiagonal_wm_pcr.cu:886-1018, 1160-1282; PCR-half-pThomas
//   :1020-1158, 1284-1431).  Blocks of 64 rows, lane t holding row 64k + t.
//   STEPS PCR steps with strides 1, 2, ..., 2^(STEPS-1) run over a sliding
//   window (each step keeps its two latest input blocks and emits a block once
//   the next one has arrived, one block behind the step before), after which row
//   i couples only to rows i -+ 2^STEPS: 64 independent subsystems of stride 64
//   (full form), or 32 of stride 32 (half form, where lanes t and t + 32 of a
//   block hold consecutive rows of subsystem t).  Thomas runs along every
//   subsystem as the blocks come out (c', d' into c, d), then backwards.
//
// Exchanges between lanes go through ds_bpermute (the reference's register
// variants, warp shuffles) or through a per-wave LDS block (its shared-memory
// variants).  Any n >= 1: rows past n are identity rows; a_0 and c_(n-1) are
// ignored.  Wave-wide synchronisation only: each wave exits on its own.
#include <stdint.h>

#include "tri_host.hpp"
#include "tri_tile.hpp"

namespace tri {

// Row exchange between the lanes of one wave: put() the lane's NV values, fetch()
// value q of any lane, done() before the next put().  LDS: through a 64*NV block
// of the wave; otherwise ds_bpermute.
template <typename T, bool LDS, int NV> struct Xch {
    T *buf;
    int lane;
    T v[NV];
    __device__ __forceinline__ void put(const T (&x)[NV]) {
        if constexpr (LDS) {
#pragma unroll
            for (int q = 0; q < NV; ++q) buf[q * 64 + lane] = x[q];
            wave_lds_sync();
        } else {
#pragma unroll
            for (int q = 0; q < NV; ++q) v[q] = x[q];
        }
    }
    __device__ __forceinline__ T fetch(int q, int src) const {
        if constexpr (LDS) return buf[q * 64 + src];
        else return lane_read(v[q], src);
    }
    __device__ __forceinline__ void done() const {
        if constexpr (LDS) wave_lds_sync();
    }
};

template <typename T> __device__ __forceinline__ T lane_bcast(T v, int l) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    }
}

template <typename T> struct BulkIO {
    __amdgpu_buffer_rsrc_t r[4];
    int64_t base, n;
    // element of array k at row i; rows outside [0, n) read the system's first row
    __device__ __forceinline__ T load(int k, int64_t i) const {
        const int64_t e = ((uint64_t)i < (uint64_t)n) ? base + i : base;
        return bload<T>(r[k], (int)(e * (int64_t)sizeof(T)), 0);
    }
    __device__ __forceinline__ void store(int k, int64_t i, T v) const {
        if (i < n) bstore<T>(v, r[k], (int)((base + i) * (int64_t)sizeof(T)), 0);
    }
    // row i as a, b, c, d: identity past n, a_0 and c_(n-1) dropped
    __device__ __forceinline__ void fix(int64_t i, T &a, T &b, T &c, T &d) const {
        const bool in = i < n;
        a = (in && i != 0) ? a : T(0);
        b = in ? b : T(1);
        c = (in && i != n - 1) ? c : T(0);
        d = in ? d : T(0);
    }
};

// ---------------------------------------------------------------- WM-pGE
template <typename T, int ROUNDS, bool LDS, int DF>
__global__ void __launch_bounds__(256) bulk_wm_kernel(const T *a, const T *b, T *c, T *d, int64_t n,
                                                      int64_t batch, int64_t sstride, int32_t bb) {
    constexpr int B = 1 << ROUNDS;            // rows (lanes) per block
    constexpr int NB = 64 / B;                // blocks per half-chunk
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t s = (int64_t)blockIdx.x * 4 + wave;
    if (s >= batch) return;                   // whole waves; no workgroup barriers
    __shared__ T xs[4][LDS ? 3 * 2 * 64 : 1];
    Xch<T, LDS, 6> X{&xs[wave][0], lane, {}};
    BulkIO<T> io{{make_rsrc(a, bb), make_rsrc(b, bb), make_rsrc(c, bb), make_rsrc(d, bb)},
                 s * sstride, n};
    const int64_t nch = (n + 127) / 128;

    // forward: rows in flight DF chunks ahead
    T ring[DF][2][4];
#pragma unroll
    for (int k = 0; k < DF; ++k)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int q = 0; q < 4; ++q) ring[k][h][q] = io.load(q, (int64_t)k * 128 + h * 64 + lane);
    T rho = 0, sig = 0;                       // previous block's last row: x_E = rho - sig x_F
    for (int64_t k0 = 0; k0 < nch; k0 += DF) {
#pragma unroll
        for (int kk = 0; kk < DF; ++kk) {
            const int64_t k = k0 + kk;
            T cur[2][4];
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    cur[h][q] = ring[kk][h][q];
                    ring[kk][h][q] = io.load(q, (k + DF) * 128 + h * 64 + lane);
                }
            if (k >= nch) continue;           // wave-uniform
            // every row in outer form relative to its own neighbours
            T r[2], p[2], qv[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                T av = cur[h][0], bv = cur[h][1], cv = cur[h][2], dv = cur[h][3];
                io.fix(k * 128 + h * 64 + lane, av, bv, cv, dv);
                const T inv = rcp(bv);
                r[h] = dv * inv;
                p[h] = av * inv;
                qv[h] = cv * inv;
            }
            // Wang-Mou merges: blocks of 2^t lanes pair up into blocks of 2^(t+1)
#pragma unroll
            for (int t = 0; t < ROUNDS; ++t) {
                const int blk = lane >> t;
                const int pb = (blk & ~1) << t;
                const int srcA = pb + (1 << t) - 1, srcB = pb + (1 << t);
                const T v6[6] = {r[0], p[0], qv[0], r[1], p[1], qv[1]};
                T fa[6], fb[6];
                X.put(v6);
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    fa[q] = X.fetch(q, srcA);
                    fb[q] = X.fetch(q, srcB);
                }
                X.done();
                const bool inA = (blk & 1) == 0;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    T RB, PB, QB, RA, PA, QA;
                    wm_pair_coeffs(fa[3 * h], fa[3 * h + 1], fa[3 * h + 2], fb[3 * h],
                                   fb[3 * h + 1], fb[3 * h + 2], RB, PB, QB, RA, PA, QA);
                    Outer<T> o{r[h], p[h], qv[h]};
                    wm_merge_row(o, inA, RB, PB, QB, RA, PA, QA);
                    r[h] = o.r;
                    p[h] = o.p;
                    qv[h] = o.q;
                }
            }
            // partial Gaussian elimination, block after block
            T cp[2] = {T(0), T(0)}, dp[2] = {T(0), T(0)};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int be = 0; be < NB; ++be) {
                    const int F = be * B, E = F + B - 1;
                    const T rF = lane_bcast(r[h], F), pF = lane_bcast(p[h], F);
                    const T qF = lane_bcast(qv[h], F);
                    const T inv = rcp(T(1) - pF * sig);
                    const T rhoF = (rF - pF * rho) * inv, sigF = qF * inv;
                    const T dv = r[h] - p[h] * rho + p[h] * sig * rhoF;
                    const T cv = qv[h] + p[h] * sig * sigF;
                    if (lane >= F && lane <= E) {
                        dp[h] = dv;
                        cp[h] = cv;
                    }
                    rho = lane_bcast(dv, E);
                    sig = lane_bcast(cv, E);
                }
                io.store(2, k * 128 + h * 64 + lane, cp[h]);
                io.store(3, k * 128 + h * 64 + lane, dp[h]);
            }
        }
    }

    // backward: x_i = d'_i - c'_i x_(first row of the next block); each lane reads
    // back what it stored
    constexpr int DB = 2 * DF;
    T back[DB][2][2];
#pragma unroll
    for (int k = 0; k < DB; ++k)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int q = 0; q < 2; ++q)
                back[k][h][q] = io.load(2 + q, (nch - 1 - k) * 128 + h * 64 + lane);
    T xa = 0;                                 // x of the row after the current block
    for (int64_t k0 = nch - 1; k0 >= 0; k0 -= DB) {
#pragma unroll
        for (int kk = 0; kk < DB; ++kk) {
            const int64_t k = k0 - kk;
            T cur[2][2];
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    cur[h][q] = back[kk][h][q];
                    back[kk][h][q] = io.load(2 + q, (k - DB) * 128 + h * 64 + lane);
                }
            if (k < 0) continue;
#pragma unroll
            for (int h = 1; h >= 0; --h) {
                const int64_t i = k * 128 + h * 64 + lane;
                const T cv = i < n ? cur[h][0] : T(0), dv = i < n ? cur[h][1] : T(0);
                T x = 0;
#pragma unroll
                for (int be = NB - 1; be >= 0; --be) {
                    const int F = be * B;
                    const T xv = dv - cv * xa;
                    if (lane >= F && lane < F + B) x = xv;
                    xa = lane_bcast(xv, F);
                }
                io.store(3, i, x);
            }
        }
    }
}

// ---------------------------------------------------------------- PCR-pThomas
template <typename T> struct Row4v {
    T a, b, c, d;
};

// One PCR step with stride s on block `cur` (lane j = row j of the block), its
// lower neighbours in `prev` and upper ones in `next` at the same level.
template <typename T, bool LDS>
__device__ __forceinline__ Row4v<T> bulk_pcr_step(Xch<T, LDS, 12> &X, const Row4v<T> &prev,
                                                  const Row4v<T> &cur, const Row4v<T> &next, int s,
                                                  int lane) {
    const int lo = (lane - s) & 63, hi = (lane + s) & 63;
    const bool fromPrev = lane < s, fromNext = lane + s >= 64;
    const T v12[12] = {prev.a, prev.b, prev.c, prev.d, cur.a, cur.b,
                       cur.c,  cur.d,  next.a, next.b, next.c, next.d};
    X.put(v12);
    // lower neighbour: prev (q 0-3) or cur (4-7) at lane lo; upper: cur or next (8-11)
    const int ql = fromPrev ? 0 : 4, qh = fromNext ? 8 : 4;
    T al, bl, cl, dl, ah, bh, ch, dh;
    if constexpr (LDS) {
        al = X.fetch(ql, lo); bl = X.fetch(ql + 1, lo); cl = X.fetch(ql + 2, lo); dl = X.fetch(ql + 3, lo);
        ah = X.fetch(qh, hi); bh = X.fetch(qh + 1, hi); ch = X.fetch(qh + 2, hi); dh = X.fetch(qh + 3, hi);
    } else {
        // ds_bpermute reads one register of the source lane: fetch both candidates
        const T al2 = X.fetch(0, lo), al3 = X.fetch(4, lo);
        const T bl2 = X.fetch(1, lo), bl3 = X.fetch(5, lo);
        const T cl2 = X.fetch(2, lo), cl3 = X.fetch(6, lo);
        const T dl2 = X.fetch(3, lo), dl3 = X.fetch(7, lo);
        const T ah2 = X.fetch(4, hi), ah3 = X.fetch(8, hi);
        const T bh2 = X.fetch(5, hi), bh3 = X.fetch(9, hi);
        const T ch2 = X.fetch(6, hi), ch3 = X.fetch(10, hi);
        const T dh2 = X.fetch(7, hi), dh3 = X.fetch(11, hi);
        al = fromPrev ? al2 : al3; bl = fromPrev ? bl2 : bl3;
        cl = fromPrev ? cl2 : cl3; dl = fromPrev ? dl2 : dl3;
        ah = fromNext ? ah3 : ah2; bh = fromNext ? bh3 : bh2;
        ch = fromNext ? ch3 : ch2; dh = fromNext ? dh3 : dh2;
    }
    X.done();
    const T k1 = cur.a * rcp(bl), k2 = cur.c * rcp(bh);
    return {-(k1 * al), cur.b - k1 * cl - k2 * ah, -(k2 * ch), cur.d - k1 * dl - k2 * dh};
}

template <typename T, int STEPS, bool LDS, int DF>
__global__ void __launch_bounds__(256) bulk_pcr_kernel(const T *a, const T *b, T *c, T *d, int64_t n,
                                                       int64_t batch, int64_t sstride, int32_t bb) {
    static_assert(STEPS == 5 || STEPS == 6, "64 or 32 subsystems");
    constexpr bool HALF = STEPS == 5;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t s = (int64_t)blockIdx.x * 4 + wave;
    if (s >= batch) return;
    __shared__ T xs[4][LDS ? 12 * 64 : 1];
    Xch<T, LDS, 12> X{&xs[wave][0], lane, {}};
    BulkIO<T> io{{make_rsrc(a, bb), make_rsrc(b, bb), make_rsrc(c, bb), make_rsrc(d, bb)},
                 s * sstride, n};
    const int64_t nb = (n + 63) / 64;
    const Row4v<T> id{T(0), T(1), T(0), T(0)};

    T ring[DF][4];
#pragma unroll
    for (int k = 0; k < DF; ++k)
#pragma unroll
        for (int q = 0; q < 4; ++q) ring[k][q] = io.load(q, (int64_t)k * 64 + lane);
    // block t of the input, DF blocks of loads in flight
    auto input = [&](int64_t t, int slot) -> Row4v<T> {
        Row4v<T> v{ring[slot][0], ring[slot][1], ring[slot][2], ring[slot][3]};
#pragma unroll
        for (int q = 0; q < 4; ++q) ring[slot][q] = io.load(q, (t + DF) * 64 + lane);
        if (t >= nb) return id;
        io.fix(t * 64 + lane, v.a, v.b, v.c, v.d);
        return v;
    };
    // blocks before the first are identity rows: every step starts with two of them
    Row4v<T> prev[STEPS], cur[STEPS];
#pragma unroll
    for (int r = 0; r < STEPS; ++r) prev[r] = cur[r] = id;
    T lc = 0, ld = 0;                          // Thomas state of the lane's subsystem row
    // at input block t, step r emits block t - r - 1
    for (int64_t t0 = 0; t0 < nb + STEPS; t0 += DF) {
#pragma unroll
        for (int kk = 0; kk < DF; ++kk) {
            const int64_t t = t0 + kk;
            Row4v<T> v = input(t, kk);
            if (t >= nb + STEPS) continue;   // wave-uniform
#pragma unroll
            for (int r = 0; r < STEPS; ++r) {
                const Row4v<T> o = bulk_pcr_step<T, LDS>(X, prev[r], cur[r], v, 1 << r, lane);
                prev[r] = cur[r];
                cur[r] = v;
                v = o;
            }
            if (t < STEPS) continue;         // blocks before the first
            // v = block t - STEPS: Thomas along the subsystems
            const int64_t i = (t - STEPS) * 64 + lane;
            if constexpr (!HALF) {
                const T piv = rcp(v.b - lc * v.a);
                lc = v.c * piv;
                ld = (v.d - ld * v.a) * piv;
            } else {
                // lanes t and t + 32 hold consecutive rows of subsystem t mod 32
                T pc = lane_read(lc, lane ^ 32), pd = lane_read(ld, lane ^ 32);
                if (lane < 32) {
                    const T piv = rcp(v.b - pc * v.a);
                    lc = v.c * piv;
                    ld = (v.d - pd * v.a) * piv;
                }
                pc = lane_read(lc, lane ^ 32);
                pd = lane_read(ld, lane ^ 32);
                if (lane >= 32) {
                    const T piv = rcp(v.b - pc * v.a);
                    lc = v.c * piv;
                    ld = (v.d - pd * v.a) * piv;
                }
            }
            io.store(2, i, lc);
            io.store(3, i, ld);
        }
    }

    // backward along every subsystem, last block first
    constexpr int DB = 2 * DF;
    T back[DB][2];
#pragma unroll
    for (int k = 0; k < DB; ++k)
#pragma unroll
        for (int q = 0; q < 2; ++q) back[k][q] = io.load(2 + q, (nb - 1 - k) * 64 + lane);
    T x = 0;                                   // x of the lane's subsystem, next row
    for (int64_t k0 = nb - 1; k0 >= 0; k0 -= DB) {
#pragma unroll
        for (int kk = 0; kk < DB; ++kk) {
            const int64_t k = k0 - kk;
            T cv = back[kk][0], dv = back[kk][1];
#pragma unroll
            for (int q = 0; q < 2; ++q) back[kk][q] = io.load(2 + q, (k - DB) * 64 + lane);
            if (k < 0) continue;
            const int64_t i = k * 64 + lane;
            if (i >= n) { cv = 0; dv = 0; }
            if constexpr (!HALF) {
                x = dv - cv * x;
            } else {
                T xn = lane_read(x, lane ^ 32);    // row after: lane t of the next block
                if (lane >= 32) x = dv - cv * xn;
                xn = lane_read(x, lane ^ 32);      // row after: lane t + 32 of this block
                if (lane < 32) x = dv - cv * xn;
            }
            io.store(3, i, x);
        }
    }
}

// ---------------------------------------------------------------- launchers
template <typename T, class K>
static hipError_t bulk_launch(K kernel, const T *a, const T *b, T *c, T *d, int64_t n,
                              int64_t batch, int64_t sstride, int32_t bb, hipStream_t st) {
    const int64_t wg = (batch + 3) / 4;
    if (wg > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kernel, dim3((unsigned)wg), dim3(256), 0, st, a, b, c, d, n, batch, sstride,
                       bb);
    return hipGetLastError();
}

// form: 0 PCR-pThomas, 1 PCR-half-pThomas, 2 WM-pGE, 3 WM-half-pGE; lds: exchanges
// through LDS (the reference's shared-memory variants) or ds_bpermute.
template <typename T>
static hipError_t launch_bulk(int form, bool lds, const T *a, const T *b, T *c, T *d, int64_t n,
                              int64_t batch, int64_t sstride, int32_t bb, hipStream_t st) {
    if (n < 1 || batch < 0 || bb <= 0 || sstride < n) return hipErrorInvalidValue;
    if (batch == 0) return hipSuccess;
#define TRI_BULK(KERNEL) return bulk_launch<T>(KERNEL, a, b, c, d, n, batch, sstride, bb, st)
    switch (form * 2 + (lds ? 1 : 0)) {
    case 0: TRI_BULK((bulk_pcr_kernel<T, 6, false, 4>));
    case 1: TRI_BULK((bulk_pcr_kernel<T, 6, true, 4>));
    case 2: TRI_BULK((bulk_pcr_kernel<T, 5, false, 4>));
    case 3: TRI_BULK((bulk_pcr_kernel<T, 5, true, 4>));
    case 4: TRI_BULK((bulk_wm_kernel<T, 6, false, 4>));
    case 5: TRI_BULK((bulk_wm_kernel<T, 6, true, 4>));
    case 6: TRI_BULK((bulk_wm_kernel<T, 5, false, 4>));
    case 7: TRI_BULK((bulk_wm_kernel<T, 5, true, 4>));
    default: return hipErrorInvalidValue;
    }
#undef TRI_BULK
}

hipError_t launch_bulk_f32(int form, bool lds, const float *a, const float *b, float *c, float *d,
                           int64_t n, int64_t batch, int64_t sstride, int32_t bb, hipStream_t st) {
    return launch_bulk<float>(form, lds, a, b, c, d, n, batch, sstride, bb, st);
}
hipError_t launch_bulk_f64(int form, bool lds, const double *a, const double *b, double *c,
                           double *d, int64_t n, int64_t batch, int64_t sstride, int32_t bb,
                           hipStream_t st) {
    return launch_bulk<double>(form, lds, a, b, c, d, n, batch, sstride, bb, st);
}

}  // namespace tri
