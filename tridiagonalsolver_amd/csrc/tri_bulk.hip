// Bulk-row forms of the reference's streaming solvers: one wave streams one
// system through its 64 lanes in row blocks, with the reference's memory
// semantics (c' into c and d' into d on the way down, x into d on the way up:
// 9 words per unknown).  These are the algorithms of the reference's ids 3-10
// themselves, for callers that want them rather than the partition kernel that
// tridiag_solve runs for those ids by default.  At the end: id 2, cyclic
// reduction with its levels in global memory (one wave per system), and ids
// 11-13, cyclic reduction over R rows per lane (one workgroup per system).
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

// ---------------------------------------------------------------- CR through global memory
// Cyclic reduction with its levels in global memory (Tridiagonal_CUDA/
// tridiagonal_basic.cu:32-164, the reference's CR id): one wave per system.
// Level l + 1 keeps the odd rows j = 2i + 1 of level l and eliminates the even
// rows around them,
//     k1 = a_j / b_(j-1),  k2 = c_j / b_(j+1),
//     a'_i = -a_(j-1) k1,  c'_i = -c_(j+1) k2,
//     b'_i = b_j - c_(j-1) k1 - a_(j+1) k2,  d'_i = d_j - d_(j-1) k1 - d_(j+1) k2,
// (row j + 1 taken as an identity row past the level's end), until at most 64
// rows are left.  Those are solved by PCR across the wave's lanes (strides 1 to
// 32), and the way back fills in the even rows of every level,
//     x_j = (d_j - a_j x_(j-1) - c_j x_(j+1)) / b_j,
// the odd rows taking x of the level above.  Levels 1 and up sit in a scratch
// buffer of 4n words per system (for level l >= 1, rows at offset
// n/2 + n/4 + ... of the l - 1 levels before it in each of four n-word arrays);
// x of a level replaces its d.  a, b and c keep their values; x replaces d.
template <typename T> struct CrRow {
    T a, b, c, d;
};

template <typename T> struct CrLevels {
    BulkIO<T> io;
    T *sa, *sb, *sc, *sd;       // this system's scratch arrays (n words each)
    // row i of the level at scratch offset `off` (level 0: the input, with
    // a_0 and c_(n-1) dropped); rows outside [0, len) are identity rows
    __device__ __forceinline__ CrRow<T> row(int lev, int64_t off, int64_t len, int64_t i) const {
        CrRow<T> r{T(0), T(1), T(0), T(0)};
        if (i < 0 || i >= len) return r;
        if (lev == 0) {
            r.a = io.load(0, i);
            r.b = io.load(1, i);
            r.c = io.load(2, i);
            r.d = io.load(3, i);
            io.fix(i, r.a, r.b, r.c, r.d);
        } else {
            r.a = sa[off + i];
            r.b = sb[off + i];
            r.c = sc[off + i];
            r.d = sd[off + i];
        }
        return r;
    }
    __device__ __forceinline__ T x(int64_t off, int64_t i) const { return sd[off + i]; }
    __device__ __forceinline__ void put_x(int lev, int64_t off, int64_t i, T v) const {
        if (lev == 0) io.store(3, i, v);
        else sd[off + i] = v;
    }
};

// Orders the global-memory writes of the wave's lanes before their later reads.
__device__ __forceinline__ void wave_global_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// rows and scratch offset of level `lev` (level 0: n rows, the input)
__device__ __forceinline__ void cr_level(int64_t n, int lev, int64_t &len, int64_t &off) {
    len = n;
    off = 0;
    for (int l = 1; l <= lev; ++l) {
        if (l >= 2) off += len;
        len /= 2;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) bulk_cr_kernel(const T *a, const T *b, T *c, T *d,
                                                      int64_t n, int64_t batch, int64_t sstride,
                                                      int32_t bb, T *scratch) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t s = (int64_t)blockIdx.x * 4 + wave;
    if (s >= batch) return;                   // whole waves; no workgroup barriers
    T *const sys = scratch + s * 4 * n;
    const CrLevels<T> L{{{make_rsrc(a, bb), make_rsrc(b, bb), make_rsrc(c, bb), make_rsrc(d, bb)},
                         s * sstride, n},
                        sys, sys + n, sys + 2 * n, sys + 3 * n};

    // forward: reduce to at most 64 rows
    int lev = 0;
    int64_t len = n, off = 0;
    while (len > 64) {
        const int64_t nlen = len / 2, noff = lev == 0 ? 0 : off + len;
        for (int64_t i0 = 0; i0 < nlen; i0 += 64) {
            const int64_t i = i0 + lane;
            if (i < nlen) {
                const int64_t j = 2 * i + 1;
                const CrRow<T> p = L.row(lev, off, len, j - 1), m = L.row(lev, off, len, j),
                               q = L.row(lev, off, len, j + 1);
                const T k1 = m.a * rcp(p.b), k2 = m.c * rcp(q.b);
                L.sa[noff + i] = -p.a * k1;
                L.sb[noff + i] = m.b - p.c * k1 - q.a * k2;
                L.sc[noff + i] = -q.c * k2;
                L.sd[noff + i] = m.d - p.d * k1 - q.d * k2;
            }
        }
        wave_global_sync();
        len = nlen;
        off = noff;
        ++lev;
    }

    // the last level (at most 64 rows, one per lane): PCR across the wave
    CrRow<T> r = L.row(lev, off, len, lane);
    for (int st = 1; st < 64; st <<= 1) {
        // every lane reads (ds_bpermute takes its data from active lanes only),
        // then rows outside the wave become identity rows
        const bool hp = lane >= st, hq = lane + st < 64;
        const int sp = hp ? lane - st : lane, sq = hq ? lane + st : lane;
        CrRow<T> p{lane_read(r.a, sp), lane_read(r.b, sp), lane_read(r.c, sp), lane_read(r.d, sp)};
        CrRow<T> q{lane_read(r.a, sq), lane_read(r.b, sq), lane_read(r.c, sq), lane_read(r.d, sq)};
        if (!hp) p = CrRow<T>{T(0), T(1), T(0), T(0)};
        if (!hq) q = CrRow<T>{T(0), T(1), T(0), T(0)};
        const T k1 = r.a * rcp(p.b), k2 = r.c * rcp(q.b);
        r = CrRow<T>{-p.a * k1, r.b - p.c * k1 - q.a * k2, -q.c * k2, r.d - p.d * k1 - q.d * k2};
    }
    if (lane < len) L.put_x(lev, off, lane, r.d * rcp(r.b));
    wave_global_sync();

    // back: the even rows of every level from the odd rows' x one level up
    for (int l = lev - 1; l >= 0; --l) {
        int64_t ln, lo, un, uo;
        cr_level(n, l, ln, lo);
        cr_level(n, l + 1, un, uo);
        const int64_t ne = (ln + 1) / 2;
        for (int64_t i0 = 0; i0 < ne; i0 += 64) {
            const int64_t i = i0 + lane;
            if (i < ne) {
                const int64_t j = 2 * i;
                const T xl = i >= 1 ? L.x(uo, i - 1) : T(0);
                const T xr = j + 1 < ln ? L.x(uo, i) : T(0);
                const CrRow<T> m = L.row(l, lo, ln, j);
                L.put_x(l, lo, j, (m.d - m.a * xl - m.c * xr) * rcp(m.b));
                if (j + 1 < ln) L.put_x(l, lo, j + 1, xr);
            }
        }
        wave_global_sync();
    }
}

// ---------------------------------------------------------------- CR over R rows per lane
// The reference's register CR (Tridiagonal_CUDA/tridiagonal_onchip.cu:170-369,
// CR_reg4/8/16): one workgroup per system, lane t holding rows tR .. tR + R - 1 in
// registers.  log2 R cyclic-reduction levels run inside the lanes: at level k
// (h = 2^k) the rows r with (r + 1) % 2h == 0 take in their neighbours r - h and
// r + h (the next lane's row h - 1 for the lane's last row, through LDS), so that
// each lane's last row couples only to the last rows of the lanes beside it.
// That system of one row per lane is solved by PCR in LDS (the reference runs CR
// there), and the eliminated rows are filled in level by level on the way back,
// the previous lane's last x standing in for row -1.  Any n up to 1024 R (the
// reference's bound); rows past n are identity rows.  a, b, c keep their values;
// x replaces d.
template <typename T, int R>
__global__ void __launch_bounds__(1024) bulk_crreg_kernel(const T *a, const T *b, T *c, T *d,
                                                          int64_t n, int64_t batch,
                                                          int64_t sstride, int32_t bb) {
    static_assert(R >= 2 && (R & (R - 1)) == 0, "R is a power of two");
    constexpr int LR = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : R == 16 ? 4 : 5;
    __shared__ T sh[4][1024];
    const int t = threadIdx.x, P = blockDim.x;
    const int64_t s = blockIdx.x;
    const BulkIO<T> io{{make_rsrc(a, bb), make_rsrc(b, bb), make_rsrc(c, bb), make_rsrc(d, bb)},
                       s * sstride, n};
    T ra[R], rb[R], rc[R], rd[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t i = (int64_t)t * R + r;
        ra[r] = io.load(0, i);
        rb[r] = io.load(1, i);
        rc[r] = io.load(2, i);
        rd[r] = io.load(3, i);
        io.fix(i, ra[r], rb[r], rc[r], rd[r]);
    }

    // reduction inside the lanes
#pragma unroll
    for (int k = 0; k < LR; ++k) {
        const int h = 1 << k;
        __syncthreads();
        sh[0][t] = ra[h - 1];
        sh[1][t] = rb[h - 1];
        sh[2][t] = rc[h - 1];
        sh[3][t] = rd[h - 1];
        __syncthreads();
        const bool last = t == P - 1;
        const T na = last ? T(0) : sh[0][t + 1], nb = last ? T(1) : sh[1][t + 1];
        const T nc = last ? T(0) : sh[2][t + 1], nd = last ? T(0) : sh[3][t + 1];
#pragma unroll
        for (int r = 2 * h - 1; r < R; r += 2 * h) {
            const int p = r - h;
            const T qa = r + h < R ? ra[r + h < R ? r + h : 0] : na;
            const T qb = r + h < R ? rb[r + h < R ? r + h : 0] : nb;
            const T qc = r + h < R ? rc[r + h < R ? r + h : 0] : nc;
            const T qd = r + h < R ? rd[r + h < R ? r + h : 0] : nd;
            const T k1 = ra[r] * rcp(rb[p]), k2 = rc[r] * rcp(qb);
            ra[r] = -ra[p] * k1;
            rb[r] = rb[r] - rc[p] * k1 - qa * k2;
            rc[r] = -qc * k2;
            rd[r] = rd[r] - rd[p] * k1 - qd * k2;
        }
    }

    // the lanes' last rows: PCR across the workgroup in LDS
    T ia = ra[R - 1], ib = rb[R - 1], ic = rc[R - 1], id = rd[R - 1];
    for (int st = 1; st < P; st <<= 1) {
        __syncthreads();
        sh[0][t] = ia;
        sh[1][t] = ib;
        sh[2][t] = ic;
        sh[3][t] = id;
        __syncthreads();
        const bool hp = t >= st, hq = t + st < P;
        const T pa = hp ? sh[0][t - st] : T(0), pb = hp ? sh[1][t - st] : T(1);
        const T pc = hp ? sh[2][t - st] : T(0), pd = hp ? sh[3][t - st] : T(0);
        const T qa = hq ? sh[0][t + st] : T(0), qb = hq ? sh[1][t + st] : T(1);
        const T qc = hq ? sh[2][t + st] : T(0), qd = hq ? sh[3][t + st] : T(0);
        const T k1 = ia * rcp(pb), k2 = ic * rcp(qb);
        ia = -pa * k1;
        ib = ib - pc * k1 - qa * k2;
        ic = -qc * k2;
        id = id - pd * k1 - qd * k2;
    }
    rd[R - 1] = id * rcp(ib);
    __syncthreads();
    sh[0][t] = rd[R - 1];
    __syncthreads();
    const T xprev = t > 0 ? sh[0][t - 1] : T(0);

    // back inside the lanes: the rows eliminated at level k, last level first
#pragma unroll
    for (int k = LR - 1; k >= 0; --k) {
        const int h = 1 << k;
#pragma unroll
        for (int r = h - 1; r < R; r += 2 * h) {
            const T xl = r >= h ? rd[r >= h ? r - h : 0] : xprev;
            rd[r] = (rd[r] - ra[r] * xl - rc[r] * rd[r + h]) * rcp(rb[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) io.store(3, (int64_t)t * R + r, rd[r]);
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

// CR through global memory; scratch: 4 * n * batch words.
template <typename T>
static hipError_t launch_bulk_cr(const T *a, const T *b, T *c, T *d, int64_t n, int64_t batch,
                                 int64_t sstride, int32_t bb, T *scratch, hipStream_t st) {
    if (n < 1 || batch < 0 || bb <= 0 || sstride < n || !scratch) return hipErrorInvalidValue;
    if (batch == 0) return hipSuccess;
    const int64_t wg = (batch + 3) / 4;
    if (wg > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((bulk_cr_kernel<T>), dim3((unsigned)wg), dim3(256), 0, st, a, b, c, d, n,
                       batch, sstride, bb, scratch);
    return hipGetLastError();
}
hipError_t launch_bulk_cr_f32(const float *a, const float *b, float *c, float *d, int64_t n,
                              int64_t batch, int64_t sstride, int32_t bb, float *scratch,
                              hipStream_t st) {
    return launch_bulk_cr<float>(a, b, c, d, n, batch, sstride, bb, scratch, st);
}
hipError_t launch_bulk_cr_f64(const double *a, const double *b, double *c, double *d, int64_t n,
                              int64_t batch, int64_t sstride, int32_t bb, double *scratch,
                              hipStream_t st) {
    return launch_bulk_cr<double>(a, b, c, d, n, batch, sstride, bb, scratch, st);
}

// Register CR over R = 4, 8 or 16 rows per lane, one workgroup per system of at
// most 1024 R rows.
template <typename T>
static hipError_t launch_bulk_crreg(int R, const T *a, const T *b, T *c, T *d, int64_t n,
                                    int64_t batch, int64_t sstride, int32_t bb, hipStream_t st) {
    if (n < 1 || batch < 0 || bb <= 0 || sstride < n || n > 1024 * (int64_t)R)
        return hipErrorInvalidValue;
    if (batch == 0) return hipSuccess;
    if (batch > 0x7fffffff) return hipErrorInvalidValue;
    const int64_t lanes = (n + R - 1) / R;
    const unsigned P = (unsigned)((lanes + 63) / 64 * 64);
    switch (R) {
    case 4:
        hipLaunchKernelGGL((bulk_crreg_kernel<T, 4>), dim3((unsigned)batch), dim3(P), 0, st, a, b, c,
                           d, n, batch, sstride, bb);
        break;
    case 8:
        hipLaunchKernelGGL((bulk_crreg_kernel<T, 8>), dim3((unsigned)batch), dim3(P), 0, st, a, b, c,
                           d, n, batch, sstride, bb);
        break;
    case 16:
        hipLaunchKernelGGL((bulk_crreg_kernel<T, 16>), dim3((unsigned)batch), dim3(P), 0, st, a, b,
                           c, d, n, batch, sstride, bb);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_bulk_crreg_f32(int R, const float *a, const float *b, float *c, float *d,
                                 int64_t n, int64_t batch, int64_t sstride, int32_t bb,
                                 hipStream_t st) {
    return launch_bulk_crreg<float>(R, a, b, c, d, n, batch, sstride, bb, st);
}
hipError_t launch_bulk_crreg_f64(int R, const double *a, const double *b, double *c, double *d,
                                 int64_t n, int64_t batch, int64_t sstride, int32_t bb,
                                 hipStream_t st) {
    return launch_bulk_crreg<double>(R, a, b, c, d, n, batch, sstride, bb, st);
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
