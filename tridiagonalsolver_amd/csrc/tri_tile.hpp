// On-chip partition solver ("tile" kernels) for gfx950.
//
// A workgroup owns S whole systems.  Every system is cut into P segments of M
// consecutive rows; one lane owns one segment and keeps it in VGPRs from the
// moment it is loaded until its x is stored, so every unknown crosses HBM
// exactly 5 times (read a, b, c, d; write x).
//
//  1. lane sweep  : a Thomas-like forward sweep over the lane's rows writes each
//                   interior row as x_j = de_j - al_j x_L - ga_j x_{j+1}
//                   (x_L = the segment's first row) and leaves the segment's
//                   first and last rows coupled only to the rows just outside
//                   the segment (Wang's partition method, H.H. Wang 1981;
//                   the per-thread form of Laszlo, Giles, Appleyard, ACM TOMS
//                   42(4) 2016).
//  2. wave merge  : log2(segments per wave) Wang-Mou merge rounds (Wang & Mou
//                   1991) across lanes make the first/last row of every segment
//                   depend only on the rows just outside the wave-block.
//  3. WG chain    : a partial Gaussian elimination (pGE) chain over the
//                   wave-blocks of a system, in LDS, gives the x values at the
//                   wave-block edges.
//  4. substitution: every lane gets x_L, x_R of its segment and back-substitutes
//                   its interior rows from registers.
//
// The reference solves the same families with one warp streaming through a
// system (Tridiagonal_CUDA/tridiagonal_wm_pcr.cu:17-205 WM-pGE,
// :886-1018 PCR-pThomas), which writes c', d' to global memory and reads them
// back (9 words per unknown); this kernel keeps them on chip.
//
// Mappings (IL = interleaved layout, element (s, i) at i*batch + s):
//   IL : lane -> system lane % S, segment (wave, lane / S).  A wave reads
//        S consecutive systems per row: S*sizeof(T) = 32 contiguous bytes,
//        and the 128-byte line is shared by 128/(32) workgroups placed on the
//        same XCD (xcd_group below) so the line is read from HBM once.
//   !IL: contiguous per system (layout 0, or layout 1 with system stride 4n);
//        a system is owned by NWS consecutive waves, lane = segment.
#pragma once
#include "tri_coop_sched.hpp"
#include "tri_device.hpp"
#include "tri_iface.hpp"

namespace tri {

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  Groups
// of G consecutive tile indices (which share 128-byte lines in the interleaved
// layout) are sent to one XCD in consecutive dispatch slots so that their L2
// holds the shared lines.
__device__ __forceinline__ int xcd_group(int b, int nblocks, int G) {
    const int span = 8 * G;
    const int nfull = (nblocks / span) * span;
    if (G <= 1 || b >= nfull) return b;
    const int xcd = b & 7, idx = b >> 3;
    const int q = idx / G, r = idx - q * G;
    return G * (8 * q + xcd) + r;
}

template <typename T> struct TileParams {
    const T *a, *b, *c;
    T *d;
    int64_t n, batch;
    int64_t sstride;  // distance between systems (IL: 1)
    int64_t rstride;  // distance between rows (IL: batch)
    // Byte extent of each coefficient array when it is below 2^31 (buffer
    // addressing with 32-bit offsets is allowed), else 0.
    int32_t buf_bytes;
    int32_t iface;     // interface solver (Iface)
    int32_t xchg_lds;  // Wang-Mou exchanges through LDS (1) or ds_bpermute (0)
    // Interleaved tiles: line groups resident at once, whose tiles are spread over the
    // batch (coop_group); 0: tiles in dispatch order.  Set by launch_tile.
    int32_t spread = 0;
};

template <typename T, int M, int NW, int S, bool IL, bool VEC, bool ALDS, bool BIG, bool PAD,
          bool NT, bool XP>
struct TileShape {
    static_assert(M >= 2, "a lane segment needs at least two rows");
    static constexpr int M_ = M;
    static constexpr int LS = IL ? S : 1;        // lane stride between segments of a system
    static constexpr int SW = 64 / LS;           // segments of one system per wave
    static constexpr int NWS = IL ? NW : NW / S; // waves per system
    static constexpr int P = SW * NWS;           // segments per system
    static constexpr int CAP = P * M;            // rows per system the tile can hold
    static constexpr int NI = (M > 2) ? (M - 2) : 1;
    static constexpr int V = (VEC && !IL) ? Vec16<T>::N : 1;      // elements per load
    // Prefetch ring: NBUF chunks of CH rows; at most (NBUF-1)*CH*4/V loads in
    // flight per wave, below the 63 outstanding loads a wave can count.
    static constexpr int CHR0 = (V > 1) ? 8 : (sizeof(T) == 4 ? 4 : 2);
    static constexpr int CHR = (M % CHR0 == 0) ? CHR0 : 4;       // M = 12: chunks of 4 rows
    static constexpr int CH = (M < CHR) ? M : CHR;
    static constexpr int NC = M / CH;
    static constexpr int NBUF_MAX = (NW >= 16) ? 2 : 3;   // 1024-lane groups: 128 VGPRs
    static constexpr int NBUF = (NC >= NBUF_MAX) ? NBUF_MAX : NC;
    static_assert(IL ? (64 % S == 0) : (NW % S == 0), "bad tile shape");
    static_assert(M % CH == 0 && CH % V == 0, "chunking must divide the segment");
};

// ---------------------------------------------------------------- loaders
// Fast path: every row of every lane lies inside the system (n == CAP) and each
// array spans < 2^31 bytes.  Buffer loads address a row as one per-lane VGPR
// offset plus a wave-uniform SGPR row offset, so in-flight loads cost no
// address registers.
// Cache-policy operand of the buffer instructions: 2 = nontemporal (nt).
template <bool NT> constexpr int cpol() { return NT ? 2 : 0; }

template <typename T, bool NT = false>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, cpol<NT>()));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, cpol<NT>()));
}
template <typename T, bool NT = false>
__device__ __forceinline__ void bstore(T x, __amdgpu_buffer_rsrc_t r, int vo, int so) {
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, vo, so,
                                              cpol<NT>());
    else
        __builtin_amdgcn_raw_buffer_store_b64(
            __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, x), r, vo, so,
            cpol<NT>());
}
using u32x4 = __attribute__((ext_vector_type(4))) unsigned;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, int32_t bytes) {
    // raw buffer (stride 0); word 3 = 0x00020000 selects 32-bit data format on gfx9
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, bytes, 0x00020000);
}

// 16-byte chunk at row e of the n-row system that starts at element `base` of the
// array behind r (a buffer below 2 GiB): the whole chunk inside the system, its
// leading rows when it straddles n, zero past n.  Never touches rows at or past n.
template <typename T, bool NT = false>
__device__ __forceinline__ u32x4 sys_chunk_load(__amdgpu_buffer_rsrc_t r, int64_t base, int64_t e,
                                                int64_t n) {
    constexpr int VV = 16 / (int)sizeof(T);
    u32x4 t = {0u, 0u, 0u, 0u};
    if (e + VV <= n) {
        t = __builtin_amdgcn_raw_buffer_load_b128(r, (int)((base + e) * (int64_t)sizeof(T)), 0,
                                                  cpol<NT>());
    } else if (e < n) {
#pragma unroll
        for (int v = 0; v < VV; ++v)
            if (e + v < n)
                reinterpret_cast<T *>(&t)[v] =
                    bload<T, NT>(r, (int)((base + e + v) * (int64_t)sizeof(T)), 0);
    }
    return t;
}
// Stores the rows of a 16-byte chunk that lie inside the system (see sys_chunk_load).
template <typename T, bool NT = false>
__device__ __forceinline__ void sys_chunk_store(u32x4 t, __amdgpu_buffer_rsrc_t r, int64_t base,
                                                int64_t e, int64_t n) {
    constexpr int VV = 16 / (int)sizeof(T);
    if (e + VV <= n) {
        __builtin_amdgcn_raw_buffer_store_b128(t, r, (int)((base + e) * (int64_t)sizeof(T)), 0,
                                               cpol<NT>());
    } else if (e < n) {
#pragma unroll
        for (int v = 0; v < VV; ++v)
            if (e + v < n)
                bstore<T, NT>(reinterpret_cast<const T *>(&t)[v], r,
                              (int)((base + e + v) * (int64_t)sizeof(T)), 0);
    }
}

// The same for a buffer that starts at the first row of the range (base 0).
template <typename T, bool NT>
__device__ __forceinline__ u32x4 tail_chunk_load(__amdgpu_buffer_rsrc_t r, int64_t e, int64_t n) {
    return sys_chunk_load<T, NT>(r, 0, e, n);
}
template <typename T, bool NT>
__device__ __forceinline__ void tail_chunk_store(u32x4 t, __amdgpu_buffer_rsrc_t r, int64_t e,
                                                 int64_t n) {
    sys_chunk_store<T, NT>(t, r, 0, e, n);
}

__device__ __forceinline__ int64_t wave_uniform(int64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

template <typename T, int CH, int V, bool NT> struct FastLoader {
    __amdgpu_buffer_rsrc_t ra, rb, rc, rd;
    int vo;       // lane byte offset of its first row
    int rowb;     // bytes between consecutive rows of one system
    __device__ __forceinline__ void load(T (&q)[4][CH], int j0) const {
        if constexpr (V == 1) {
#pragma unroll
            for (int v = 0; v < CH; ++v) {
                const int so = (j0 + v) * rowb;
                q[0][v] = bload<T, NT>(ra, vo, so);
                q[1][v] = bload<T, NT>(rb, vo, so);
                q[2][v] = bload<T, NT>(rc, vo, so);
                q[3][v] = bload<T, NT>(rd, vo, so);
            }
        } else {
#pragma unroll
            for (int v0 = 0; v0 < CH; v0 += V) {
                const int so = (j0 + v0) * rowb;
                u32x4 xa = __builtin_amdgcn_raw_buffer_load_b128(ra, vo, so, cpol<NT>());
                u32x4 xb = __builtin_amdgcn_raw_buffer_load_b128(rb, vo, so, cpol<NT>());
                u32x4 xc = __builtin_amdgcn_raw_buffer_load_b128(rc, vo, so, cpol<NT>());
                u32x4 xd = __builtin_amdgcn_raw_buffer_load_b128(rd, vo, so, cpol<NT>());
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    q[0][v0 + v] = reinterpret_cast<const T *>(&xa)[v];
                    q[1][v0 + v] = reinterpret_cast<const T *>(&xb)[v];
                    q[2][v0 + v] = reinterpret_cast<const T *>(&xc)[v];
                    q[3][v0 + v] = reinterpret_cast<const T *>(&xd)[v];
                }
            }
        }
    }
};

// Padded tiles (n < CAP): the same buffer addressing, rows past the end of the
// system are identity rows (0, 1, 0, 0) and are not read.
template <typename T, int CH, int V, bool NT = false> struct PredLoader {
    __amdgpu_buffer_rsrc_t ra, rb, rc, rd;
    int vo, rowb;
    int64_t row0, n;
    __device__ __forceinline__ void load(T (&q)[4][CH], int j0) const {
        if (row0 + j0 + CH <= n) {
            // the whole chunk lies inside the system (every chunk but one per system):
            // the unpredicated loads of FastLoader, without a branch per row
            FastLoader<T, CH, V, NT>{ra, rb, rc, rd, vo, rowb}.load(q, j0);
            return;
        }
#pragma unroll
        for (int v0 = 0; v0 < CH; v0 += V) {
            const int64_t r = row0 + j0 + v0;
            const int so = (j0 + v0) * rowb;
            if (V > 1 && r + V <= n) {
                u32x4 xa = __builtin_amdgcn_raw_buffer_load_b128(ra, vo, so, cpol<NT>());
                u32x4 xb = __builtin_amdgcn_raw_buffer_load_b128(rb, vo, so, cpol<NT>());
                u32x4 xc = __builtin_amdgcn_raw_buffer_load_b128(rc, vo, so, cpol<NT>());
                u32x4 xd = __builtin_amdgcn_raw_buffer_load_b128(rd, vo, so, cpol<NT>());
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    q[0][v0 + v] = reinterpret_cast<const T *>(&xa)[v];
                    q[1][v0 + v] = reinterpret_cast<const T *>(&xb)[v];
                    q[2][v0 + v] = reinterpret_cast<const T *>(&xc)[v];
                    q[3][v0 + v] = reinterpret_cast<const T *>(&xd)[v];
                }
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    if (r + v < n) {
                        const int sov = so + v * rowb;
                        q[0][v0 + v] = bload<T, NT>(ra, vo, sov);
                        q[1][v0 + v] = bload<T, NT>(rb, vo, sov);
                        q[2][v0 + v] = bload<T, NT>(rc, vo, sov);
                        q[3][v0 + v] = bload<T, NT>(rd, vo, sov);
                    } else {
                        q[0][v0 + v] = 0; q[1][v0 + v] = 1; q[2][v0 + v] = 0; q[3][v0 + v] = 0;
                    }
                }
            }
        }
    }
};

// Rows already staged in registers (transposed wave-block loads).
template <typename T, int CH, int M> struct RegLoader {
    const T (*rows)[M];
    __device__ __forceinline__ void load(T (&q)[4][CH], int j0) const {
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int v = 0; v < CH; ++v) q[k][v] = rows[k][j0 + v];
    }
};

// Arrays of 2 GiB or more: 64-bit addressing, identity rows past the end.
template <typename T, int CH, int V> struct SafeLoader {
    const T *A, *B, *C, *D;
    int64_t base, rs, row0, n;
    bool valid;
    __device__ __forceinline__ void load(T (&q)[4][CH], int j0) const {
#pragma unroll
        for (int v0 = 0; v0 < CH; v0 += V) {
            const int64_t r = row0 + j0 + v0;
            if (V > 1 && valid && r + V <= n) {
                using VT = typename Vec16<T>::type;
                const int64_t off = base + r;
                VT xa = *reinterpret_cast<const VT *>(A + off);
                VT xb = *reinterpret_cast<const VT *>(B + off);
                VT xc = *reinterpret_cast<const VT *>(C + off);
                VT xd = *reinterpret_cast<const VT *>(D + off);
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    q[0][v0 + v] = reinterpret_cast<const T *>(&xa)[v];
                    q[1][v0 + v] = reinterpret_cast<const T *>(&xb)[v];
                    q[2][v0 + v] = reinterpret_cast<const T *>(&xc)[v];
                    q[3][v0 + v] = reinterpret_cast<const T *>(&xd)[v];
                }
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const int64_t rr = r + v;
                    if (valid && rr < n) {
                        const int64_t off = base + rr * rs;
                        q[0][v0 + v] = A[off];
                        q[1][v0 + v] = B[off];
                        q[2][v0 + v] = C[off];
                        q[3][v0 + v] = D[off];
                    } else {
                        q[0][v0 + v] = 0; q[1][v0 + v] = 1; q[2][v0 + v] = 0; q[3][v0 + v] = 0;
                    }
                }
            }
        }
    }
};

// Interleaved whole-piece loads ("XP-IL"): rows [wrow, wrow + SW*M) of the S
// consecutive systems starting at col0, where every row piece is S*sizeof(T)
// contiguous bytes (CPR 16-byte chunks).  Consecutive lanes read consecutive
// chunks (il_block_issue, NQ = M*sizeof(T)/16 chunks per lane and array); lane
// (system sl, segment kw) then takes its M rows of system sl from the wave's LDS
// block xb (il_block_transpose).  Offsets are 32-bit: the arrays are below 2 GiB.
template <typename T, int M, int SW, int S, bool NT>
__device__ __forceinline__ void il_block_issue(const __amdgpu_buffer_rsrc_t (&rr)[4], int64_t wrow,
                                               int64_t rs, int64_t col0, int lane,
                                               u32x4 (&g)[4][M * (int)sizeof(T) / 16]) {
    constexpr int CPR = S * (int)sizeof(T) / 16;
    constexpr int RB = SW * M;
    constexpr int NQ = RB * CPR / 64;
    static_assert((RB * CPR) % 64 == 0 && NQ == M * (int)sizeof(T) / 16,
                  "wave block must be whole 16-byte loads");
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int cidx = q * 64 + lane, r = cidx / CPR, h = cidx % CPR;
        const int off = (int)(((wrow + r) * rs + col0) * (int64_t)sizeof(T)) + h * 16;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            g[k][q] = __builtin_amdgcn_raw_buffer_load_b128(rr[k], off, 0, cpol<NT>());
    }
}

template <typename T, int M, int S>
__device__ __forceinline__ void il_block_transpose(const u32x4 (&g)[4][M * (int)sizeof(T) / 16],
                                                   int lane, int sl, int kw, u32x4 *xb,
                                                   T (&rows)[4][M]) {
    constexpr int NQ = M * (int)sizeof(T) / 16;
    const T *tb = reinterpret_cast<const T *>(xb);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k > 0) wave_lds_sync();
#pragma unroll
        for (int q = 0; q < NQ; ++q) xb[q * 64 + lane] = g[k][q];
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < M; ++j) rows[k][j] = tb[(kw * M + j) * S + sl];
    }
}

template <typename T, int M, int SW, int S, bool NT>
__device__ __forceinline__ void il_block_load(const __amdgpu_buffer_rsrc_t (&rr)[4], int64_t wrow,
                                              int64_t rs, int64_t col0, int lane, int sl, int kw,
                                              u32x4 *xb, T (&rows)[4][M]) {
    u32x4 g[4][M * (int)sizeof(T) / 16];
    il_block_issue<T, M, SW, S, NT>(rr, wrow, rs, col0, lane, g);
    il_block_transpose<T, M, S>(g, lane, sl, kw, xb, rows);
}

// Stores x of the same block (xs[j] = x of the lane's row j) the same way; chunks
// of systems at or past `batch` are not written.  Every lane of the wave takes part.
template <typename T, int M, int SW, int S, bool NT>
__device__ __forceinline__ void il_block_store(__amdgpu_buffer_rsrc_t rd, int64_t wrow, int64_t rs,
                                               int64_t col0, int64_t batch, int lane, int sl,
                                               int kw, u32x4 *xb, const T (&xs)[M]) {
    constexpr int VV = 16 / (int)sizeof(T);
    constexpr int CPR = S * (int)sizeof(T) / 16;
    constexpr int RB = SW * M;
    constexpr int NQ = RB * CPR / 64;
    T *tb = reinterpret_cast<T *>(xb);
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < M; ++j) tb[(kw * M + j) * S + sl] = xs[j];
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int cidx = q * 64 + lane, r = cidx / CPR, h = cidx % CPR;
        const int off = (int)(((wrow + r) * rs + col0) * (int64_t)sizeof(T)) + h * 16;
        const int64_t s0 = col0 + h * VV;                  // first system of the chunk
        const u32x4 t = xb[cidx];
        if (s0 + VV <= batch) {
            __builtin_amdgcn_raw_buffer_store_b128(t, rd, off, 0, cpol<NT>());
        } else {
#pragma unroll
            for (int v = 0; v < VV; ++v)
                if (s0 + v < batch)
                    bstore<T, NT>(reinterpret_cast<const T *>(&t)[v], rd, off + v * (int)sizeof(T), 0);
        }
    }
}

// ---------------------------------------------------------------- lane sweep
// Forward modified-Thomas over the lane's interior rows j = 1..M-2 with x_L kept
// symbolic.  Leaves x_j = de_j - al_j x_L - ga_j x_{j+1} for every interior row
// and accumulates x_1 = SD - SA x_L + PI x_R on the way:
//   x_1 = sum_j pi_j (de_j - al_j x_L) + pi_{M-1} x_R,  pi_1 = 1, pi_{j+1} = -ga_j pi_j.
template <typename T, int NI> struct SweepOut {
    T ga[NI], de[NI];
    T aL, bL, cL, dL, aR, bR, cR, dR;
    T g, alp, del, SD, SA, PI;
};

template <typename T, class Sh, class Loader, class AlStore>
__device__ __forceinline__ void lane_sweep(SweepOut<T, Sh::NI> &o, const Loader &ld,
                                           int64_t row0, int64_t n, AlStore al_store) {
    constexpr int CH = Sh::CH, NC = Sh::NC, NBUF = Sh::NBUF, M = Sh::M_;
    o.aL = 0; o.bL = 1; o.cL = 0; o.dL = 0; o.aR = 0; o.bR = 1; o.cR = 0; o.dR = 0;
    o.g = 0; o.alp = -1; o.del = 0;   // virtual row before the first interior row
    o.SD = 0; o.SA = 0; o.PI = 1;
    T buf[NBUF][4][CH];
#pragma unroll
    for (int p = 0; p < NBUF - 1; ++p) ld.load(buf[p], p * CH);
#pragma unroll
    for (int ck = 0; ck < NC; ++ck) {
        if (ck + NBUF - 1 < NC) ld.load(buf[(ck + NBUF - 1) % NBUF], (ck + NBUF - 1) * CH);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int v = 0; v < CH; ++v) {
            const int j = ck * CH + v;
            const int64_t r = row0 + j;
            T a = buf[ck % NBUF][0][v], b = buf[ck % NBUF][1][v];
            T c = buf[ck % NBUF][2][v], d = buf[ck % NBUF][3][v];
            a = (r == 0) ? T(0) : a;            // a_0 is not part of A
            c = (r == n - 1) ? T(0) : c;        // nor is c_{n-1}
            if (j == 0) {
                o.aL = a; o.bL = b; o.cL = c; o.dL = d;
            } else if (j == M - 1) {
                o.aR = a; o.bR = b; o.cR = c; o.dR = d;
            } else {
                const T rr = rcp(b - a * o.g);
                o.g = c * rr;
                o.alp = -(a * o.alp) * rr;
                o.del = (d - a * o.del) * rr;
                o.SD += o.PI * o.del;
                o.SA += o.PI * o.alp;
                o.PI = -o.g * o.PI;
                al_store(j - 1, o.alp);
                o.ga[j - 1] = o.g;
                o.de[j - 1] = o.del;
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Outer form (first row oL, last row oR) of a lane segment from its reduced rows
//   L: aL u + DLr x_L + ELr x_R = FLr ;  R: ARr x_L + DRr x_R + cR v = FRr.
template <typename T>
__device__ __forceinline__ void segment_outer(T aL, T DLr, T ELr, T FLr, T ARr, T DRr, T cR,
                                              T FRr, Outer<T> &oL, Outer<T> &oR) {
    const T inv = rcp(DLr * DRr - ELr * ARr);
    oL.r = (DRr * FLr - ELr * FRr) * inv;
    oL.p = (DRr * aL) * inv;
    oL.q = -(ELr * cR) * inv;
    oR.r = (DLr * FRr - ARr * FLr) * inv;
    oR.p = -(ARr * aL) * inv;
    oR.q = (DLr * cR) * inv;
}

// Wang-Mou merges of the SW segments of a system inside one wave (segment kw of
// system slot sl sits in lane sl % LS + LS * kw) into blocks of BS segments:
// afterwards oL / oR are in outer form relative to the block's neighbours.
// lds_x exchanges through the wave's 6*64-element LDS area xs instead of
// ds_bpermute.
template <typename T, int SW, int LS>
__device__ __forceinline__ void wm_wave_merges(Outer<T> &oL, Outer<T> &oR, int lane, int sl,
                                               int kw, int BS, bool lds_x, T *xs) {
#pragma unroll
    for (int t = 0; (1 << t) < SW; ++t) {
        if ((1 << t) >= BS) break;
        const int h = 1 << t;
        const int blk = kw >> t;
        const int pb = (blk & ~1) << t;              // first segment of the merged pair
        const int srcA = sl % LS + LS * (pb + h - 1); // lane with A's last row
        const int srcB = sl % LS + LS * (pb + h);     // lane with B's first row
        T rA, pA, qA, rB, pB, qB;
        if (lds_x) {
            xs[lane * 6 + 0] = oL.r; xs[lane * 6 + 1] = oL.p; xs[lane * 6 + 2] = oL.q;
            xs[lane * 6 + 3] = oR.r; xs[lane * 6 + 4] = oR.p; xs[lane * 6 + 5] = oR.q;
            wave_lds_sync();
            rA = xs[srcA * 6 + 3]; pA = xs[srcA * 6 + 4]; qA = xs[srcA * 6 + 5];
            rB = xs[srcB * 6 + 0]; pB = xs[srcB * 6 + 1]; qB = xs[srcB * 6 + 2];
            wave_lds_sync();
        } else {
            rA = lane_read(oR.r, srcA); pA = lane_read(oR.p, srcA); qA = lane_read(oR.q, srcA);
            rB = lane_read(oL.r, srcB); pB = lane_read(oL.p, srcB); qB = lane_read(oL.q, srcB);
        }
        T RB, PB, QB, RA, PA, QA;
        wm_pair_coeffs(rA, pA, qA, rB, pB, qB, RB, PB, QB, RA, PA, QA);
        const bool inA = (blk & 1) == 0;
        wm_merge_row(oL, inA, RB, PB, QB, RA, PA, QA);
        wm_merge_row(oR, inA, RB, PB, QB, RA, PA, QA);
    }
}

// Solves the interface system of a workgroup of tiles from every lane's reduced
// rows   L: aL u + DLr x_L + ELr x_R = FLr ;  R: ARr x_L + DRr x_R + cR v = FRr
// and returns x at the lane's first (xL) and last (xR) rows.  Every lane of the
// workgroup must call it (it synchronises the workgroup).
template <typename T, int S, int SW, int NWS, int LS, int P, int LANES>
__device__ __forceinline__ void interface_solve(T aL, T DLr, T ELr, T FLr, T ARr, T DRr, T cR,
                                                T FRr, int tid, int lane, int sl, int kw, int ws,
                                                int iface, int xchg_lds, char *tri_dyn_lds,
                                                T &xL, T &xR) {
    const int seg = ws * SW + kw;            // segment index inside the system
    const int fam = iface_family(iface);
    if (fam >= kIfacePCR) {
        // PCR / PCR + Thomas / CR on the 2P interface rows, through LDS
        IfaceLds<T> IL_ = iface_lds<T>(tri_dyn_lds, S, 2 * P);
        Row4<T> rl{aL, DLr, ELr, FLr}, rr{ARr, DRr, cR, FRr};
        if (fam == kIfacePCR || fam == kIfacePCRThomas) {
            // PCR stops at segment stride lim (depth d: 2^(d-1)); lim = P is full PCR
            const int depth = iface_pcr_depth(iface);
            int lim = (fam == kIfacePCR) ? P : (depth > 0 ? (1 << (depth - 1)) : SW);
            if (lim > P / 2) lim = P;
            pcr_interface(IL_, sl, seg, P, lim, rl, rr);
            if (lim == P) {
                xL = rl.d * rcp(rl.b);
                xR = rr.d * rcp(rr.b);
            } else {
                IL_.put(sl, 2 * seg, rl);
                IL_.put(sl, 2 * seg + 1, rr);
                __syncthreads();
                strided_thomas(IL_, S, P, lim, tid);
                xL = IL_.d[sl * 2 * P + 2 * seg];
                xR = IL_.d[sl * 2 * P + 2 * seg + 1];
            }
        } else {
            IL_.put(sl, 2 * seg, rl);
            IL_.put(sl, 2 * seg + 1, rr);
            __syncthreads();
            cr_solve(IL_, S, tid, LANES);
            xL = IL_.d[sl * 2 * P + 2 * seg];
            xR = IL_.d[sl * 2 * P + 2 * seg + 1];
        }
        return;
    }
    // Wang-Mou: outer form of the segment, merges across lanes, pGE chain across blocks
    Outer<T> oL, oR;
    segment_outer(aL, DLr, ELr, FLr, ARr, DRr, cR, FRr, oL, oR);
    const int half = fam == kIfaceWMHalf ? 1 : 0;
    const int BS = SW >> half;                   // segments per merged block
    T *xs = reinterpret_cast<T *>(tri_dyn_lds) + (size_t)(tid & ~63) * 6;
    wm_wave_merges<T, SW, LS>(oL, oR, lane, sl, kw, BS, xchg_lds != 0, xs);

    // pGE chain over the NB = NWS << half merged blocks of each system
    T u = 0, v = 0;
    const int NB = NWS << half;
    if (NB > 1) {
        __shared__ T sh_blk[S][2 * NWS][6];
        __shared__ T sh_uv[S][2 * NWS][2];
        const int bi = ws * (1 << half) + kw / BS;
        if (kw % BS == 0) {
            sh_blk[sl][bi][0] = oL.r; sh_blk[sl][bi][1] = oL.p; sh_blk[sl][bi][2] = oL.q;
        }
        if (kw % BS == BS - 1) {
            sh_blk[sl][bi][3] = oR.r; sh_blk[sl][bi][4] = oR.p; sh_blk[sl][bi][5] = oR.q;
        }
        __syncthreads();
        if (tid < S) {
            const int q = tid;
            T rho = 0, sig = 0;
            // forward: x_{E_{w-1}} = rho - sig x_{F_w}
            for (int w = 0; w < NB; ++w) {
                T rF = sh_blk[q][w][0], pF = sh_blk[q][w][1], qF = sh_blk[q][w][2];
                T rE = sh_blk[q][w][3], pE = sh_blk[q][w][4], qE = sh_blk[q][w][5];
                T inv = rcp(T(1) - pF * sig);
                T rhoF = (rF - pF * rho) * inv;
                T sigF = qF * inv;
                T rhoE = rE - pE * (rho - sig * rhoF);
                T sigE = qE + pE * sig * sigF;
                sh_blk[q][w][0] = rhoF; sh_blk[q][w][1] = sigF;
                sh_blk[q][w][2] = rhoE; sh_blk[q][w][3] = sigE;
                rho = rhoE; sig = sigE;
            }
            // backward: x_{F_w} and x_{E_w} from x_{F_{w+1}}
            T xFn = 0;
            sh_uv[q][0][0] = 0;
            for (int w = NB - 1; w >= 0; --w) {
                T xE = sh_blk[q][w][2] - sh_blk[q][w][3] * xFn;
                T xF = sh_blk[q][w][0] - sh_blk[q][w][1] * xFn;
                sh_uv[q][w][1] = xFn;
                if (w + 1 < NB) sh_uv[q][w + 1][0] = xE;
                xFn = xF;
            }
        }
        __syncthreads();
        u = sh_uv[sl][bi][0];
        v = sh_uv[sl][bi][1];
    }
    xL = oL.r - oL.p * u - oL.q * v;
    xR = oR.r - oR.p * u - oR.q * v;
}

// Occupancy floor of a tile kernel: 512-lane transposed-load tiles are held to two
// workgroups per CU (4 waves per SIMD, 128 VGPRs); one workgroup of 8 waves leaves
// too few loads in flight to stream at HBM speed.
template <typename T, int M, int NW, bool XP> constexpr int tile_min_waves() {
    return (XP && NW == 8 && M * (int)sizeof(T) <= 64) ? 4 : 1;
}

// PAD = false: the system fills the tile exactly (n == CAP), loads are not
// predicated.  PAD = true: n < CAP, rows past the end are identity rows.
// BIG: 64-bit addressing for arrays of 2 GiB or more (always predicated).
// NT: nontemporal (streaming) cache policy on the coefficient loads and the x
// stores; the data is touched once per solve.
// XP (contiguous layout, exact tiles): a wave reads its 64*M-row block of each
// array with fully coalesced 16-byte loads and hands every lane its own rows
// through LDS; x goes back the same way.  Every load instruction then covers
// whole cache lines, which is what the nontemporal policy needs.
template <typename T, int M, int NW, int S, bool IL, bool VEC, bool ALDS, bool BIG, bool PAD,
          bool NT, bool XP>
__global__ void __launch_bounds__(64 * NW)
    __attribute__((amdgpu_waves_per_eu(tile_min_waves<T, M, NW, XP>()))) tile_wm_kernel(TileParams<T> prm) {
    using Sh = TileShape<T, M, NW, S, IL, VEC, ALDS, BIG, PAD, NT, XP>;
    static_assert(!XP || (!BIG && (IL ? (!PAD && (S * (int)sizeof(T)) % 16 == 0) : VEC)),
                  "transposed loads need 16-byte rows (and exact tiles when interleaved)");
    constexpr int SW = Sh::SW, NWS = Sh::NWS, LS = Sh::LS, V = Sh::V, CH = Sh::CH;
    constexpr int LANES = 64 * NW;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int sl = IL ? (lane % S) : (wave / NWS);   // system inside the workgroup
    const int kw = IL ? (lane / S) : lane;           // segment inside the wave-block
    const int ws = IL ? wave : (wave % NWS);         // wave-block inside the system
    const int64_t n = prm.n;
    int tile = blockIdx.x;
    if (IL) {
        // tiles that share 128-byte lines on one XCD, then those line groups spread
        // over the batch so that the resident ones read whole rows, not a strip
        constexpr int G = (S * (int)sizeof(T) >= 128) ? 1 : 128 / (S * (int)sizeof(T));
        tile = xcd_group(blockIdx.x, gridDim.x, G);
        if (prm.spread > 1) {
            const int lg = tile / G;
            tile = coop_group(lg, (int)gridDim.x / G, prm.spread) * G + (tile - lg * G);
        }
    }
    const int64_t sys = (int64_t)tile * S + sl;
    const bool valid = sys < prm.batch;
    const int64_t sysc = valid ? sys : prm.batch - 1;   // loads of idle lanes stay in bounds
    const int64_t row0 = (int64_t)(ws * SW + kw) * M;
    const int64_t base = sysc * prm.sstride;
    const int64_t rs = prm.rstride;
    constexpr bool fast = !BIG && !PAD;

    // Interior coefficients; al lives in LDS when ALDS (it is only read back in
    // the substitution), laid out [row][lane] so a wave's accesses are contiguous.
    __shared__ T sh_al[ALDS ? Sh::NI : 1][ALDS ? LANES : 1];
    // transposition block of a wave: contiguous tiles 64 segments x M rows; interleaved
    // tiles SW*M rows x S systems (the same element count)
    __shared__ u32x4 xbuf[XP ? NW : 1][XP ? 64 * M * (int)sizeof(T) / 16 : 1];
    (void)xbuf;
    extern __shared__ __attribute__((aligned(16))) char tri_dyn_lds[];
    T al_r[ALDS ? 1 : Sh::NI];
    (void)al_r;
    (void)sh_al;
    auto al_store = [&](int j, T x) {
        if constexpr (ALDS) sh_al[j][tid] = x;
        else al_r[j] = x;
    };

    SweepOut<T, Sh::NI> o;
    const int32_t nb = BIG ? 0x7fffffff : prm.buf_bytes;
    __amdgpu_buffer_rsrc_t rd_rsrc = make_rsrc(prm.d, nb);
    const int vo = BIG ? 0 : (int)(((IL ? row0 * rs : row0) + base) * (int64_t)sizeof(T));
    const int rowb = (int)(rs * (int64_t)sizeof(T));
    if constexpr (BIG) {
        SafeLoader<T, CH, V> ld{prm.a, prm.b, prm.c, prm.d, base, rs, row0, n, valid};
        lane_sweep<T, Sh>(o, ld, row0, n, al_store);
    } else {
        __amdgpu_buffer_rsrc_t ra = make_rsrc(prm.a, nb), rb = make_rsrc(prm.b, nb),
                               rc = make_rsrc(prm.c, nb);
        if constexpr (XP && IL) {
            // wave block: rows [ws*SW*M, (ws+1)*SW*M) of the tile's S systems, read as
            // whole row pieces and handed to the lanes through LDS
            const __amdgpu_buffer_rsrc_t rr[4] = {ra, rb, rc, rd_rsrc};
            T rows[4][M];
            il_block_load<T, M, SW, S, NT>(rr, (int64_t)ws * SW * M, rs, (int64_t)tile * S, lane,
                                           sl, kw, xbuf[wave], rows);
            RegLoader<T, CH, M> ld{rows};
            lane_sweep<T, Sh>(o, ld, row0, n, al_store);
        } else if constexpr (XP) {
            // wave block: 64 segments x M rows = 64*M contiguous elements per array
            constexpr int NQ = 64 * M * (int)sizeof(T) / 1024;   // 16-byte loads per lane
            constexpr int PL = M * (int)sizeof(T) / 16;          // 16-byte pieces per lane row set
            constexpr int VV = 16 / (int)sizeof(T);
            const int wvo = (int)((base + (int64_t)ws * 64 * M) * (int64_t)sizeof(T)) + lane * 16;
            const int64_t e0 = (int64_t)ws * 64 * M + lane * VV;   // first row of the lane's chunk
            (void)wvo;
            (void)e0;
            u32x4 g[4][NQ];
            const __amdgpu_buffer_rsrc_t rr[4] = {ra, rb, rc, rd_rsrc};
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    if constexpr (PAD) {
                        // No bounds branch: its copy would wait for the load in flight
                        // and serialise the loads.  A chunk that holds a row below n
                        // ends inside the system (the stride is a multiple of VV and at
                        // least n); chunks at or past n read the system's first chunk
                        // and become identity rows below.
                        const int64_t e = e0 + q * 64 * VV;
                        const int off = (int)((base + (e < n ? e : 0)) * (int64_t)sizeof(T));
                        g[k][q] = __builtin_amdgcn_raw_buffer_load_b128(rr[k], off, 0, cpol<NT>());
                    } else
                        g[k][q] = __builtin_amdgcn_raw_buffer_load_b128(rr[k], wvo, q * 1024,
                                                                        cpol<NT>());
                }
            T rows[4][M];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (k > 0) wave_lds_sync();       // previous array fully read by all lanes
#pragma unroll
                for (int q = 0; q < NQ; ++q) xbuf[wave][q * 64 + lane] = g[k][q];
                wave_lds_sync();                  // block written before lanes read their rows
#pragma unroll
                for (int i = 0; i < PL; ++i) {
                    u32x4 t = xbuf[wave][lane * PL + i];
#pragma unroll
                    for (int v = 0; v < 16 / (int)sizeof(T); ++v)
                        rows[k][i * (16 / (int)sizeof(T)) + v] = reinterpret_cast<const T *>(&t)[v];
                }
            }
            if constexpr (PAD) {
#pragma unroll
                for (int j = 0; j < M; ++j)   // rows past the system: identity rows
                    if (row0 + j >= n) { rows[0][j] = 0; rows[1][j] = 1; rows[2][j] = 0; rows[3][j] = 0; }
            }
            RegLoader<T, CH, M> ld{rows};
            lane_sweep<T, Sh>(o, ld, row0, n, al_store);
        } else if constexpr (fast) {
            FastLoader<T, CH, V, NT> ld{ra, rb, rc, rd_rsrc, vo, rowb};
            lane_sweep<T, Sh>(o, ld, row0, n, al_store);
        } else {
            PredLoader<T, CH, V> ld{ra, rb, rc, rd_rsrc, vo, rowb, row0, n};
            lane_sweep<T, Sh>(o, ld, row0, n, al_store);
        }
    }

    // Reduced rows of the segment:
    //   L: aL u + DLr x_L + ELr x_R = FLr ;  R: ARr x_L + DRr x_R + cR v = FRr
    const T DLr = o.bL - o.cL * o.SA, ELr = o.cL * o.PI, FLr = o.dL - o.cL * o.SD;
    const T ARr = -(o.aR * o.alp), DRr = o.bR - o.aR * o.g, FRr = o.dR - o.aR * o.del;

    // ---------------------------------------------------------------- 2-3. interface
    T xL, xR;
    interface_solve<T, S, SW, NWS, LS, Sh::P, LANES>(o.aL, DLr, ELr, FLr, ARr, DRr, o.cR, FRr, tid,
                                                     lane, sl, kw, ws, prm.iface, prm.xchg_lds,
                                                     tri_dyn_lds, xL, xR);

    // ---------------------------------------------------------------- 4. substitution
    T *de = o.de;
    if constexpr (M > 2) {
        T xn = xR;
#pragma unroll
        for (int j = M - 3; j >= 0; --j) {
            T alj;
            if constexpr (ALDS) alj = sh_al[j][tid];
            else alj = al_r[j];
            xn = (de[j] - alj * xL) - o.ga[j] * xn;
            de[j] = xn;
        }
    }
    if constexpr (XP && IL) {
        // every lane takes part: chunk stores are assigned by position, not by system
        T xs[M];
#pragma unroll
        for (int j = 0; j < M; ++j)
            xs[j] = (j == 0) ? xL : ((j == M - 1) ? xR : de[(j >= 1 && j <= M - 2) ? j - 1 : 0]);
        il_block_store<T, M, SW, S, NT>(rd_rsrc, (int64_t)ws * SW * M, rs, (int64_t)tile * S,
                                        prm.batch, lane, sl, kw, xbuf[wave], xs);
        return;
    }
    if (!valid) return;
    if constexpr (XP) {
        constexpr int NQ = 64 * M * (int)sizeof(T) / 1024;
        constexpr int PL = M * (int)sizeof(T) / 16;
        constexpr int VV = 16 / (int)sizeof(T);
        wave_lds_sync();
#pragma unroll
        for (int i = 0; i < PL; ++i) {
            u32x4 t;
#pragma unroll
            for (int v = 0; v < VV; ++v) {
                const int j = i * VV + v;
                reinterpret_cast<T *>(&t)[v] =
                    (j == 0) ? xL : ((j == M - 1) ? xR : de[(j >= 1 && j <= M - 2) ? j - 1 : 0]);
            }
            xbuf[wave][lane * PL + i] = t;
        }
        wave_lds_sync();                          // all rows written before the block is read
        const int wvo = (int)((base + (int64_t)ws * 64 * M) * (int64_t)sizeof(T)) + lane * 16;
        const int64_t e0 = (int64_t)ws * 64 * M + lane * VV;
        (void)wvo;
        (void)e0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if constexpr (PAD)
                sys_chunk_store<T, NT>(xbuf[wave][q * 64 + lane], rd_rsrc, base, e0 + q * 64 * VV, n);
            else
                __builtin_amdgcn_raw_buffer_store_b128(xbuf[wave][q * 64 + lane], rd_rsrc, wvo,
                                                       q * 1024, cpol<NT>());
        }
        return;
    }
    T *__restrict__ D = prm.d;
    if (!BIG && !fast && row0 + M <= n) {
        // padded tiles: a segment wholly inside the system (all but one per system)
        // stores without a predicate per row
#pragma unroll
        for (int j0 = 0; j0 < M; j0 += V) {
            T xv[V];
#pragma unroll
            for (int vv = 0; vv < V; ++vv) {
                const int j = j0 + vv;
                xv[vv] = (j == 0) ? xL : ((j == M - 1) ? xR : de[(j >= 1 && j <= M - 2) ? j - 1 : 0]);
            }
            if (V > 1) {
                u32x4 xs;
#pragma unroll
                for (int vv = 0; vv < V; ++vv) reinterpret_cast<T *>(&xs)[vv] = xv[vv];
                __builtin_amdgcn_raw_buffer_store_b128(xs, rd_rsrc, vo, j0 * rowb, cpol<NT>());
            } else {
                bstore<T, NT>(xv[0], rd_rsrc, vo, j0 * rowb);
            }
        }
        return;
    }
#pragma unroll
    for (int j0 = 0; j0 < M; j0 += V) {
        T xv[V];
#pragma unroll
        for (int vv = 0; vv < V; ++vv) {
            const int j = j0 + vv;
            xv[vv] = (j == 0) ? xL : ((j == M - 1) ? xR : de[(j >= 1 && j <= M - 2) ? j - 1 : 0]);
        }
        if (!BIG) {
            const int64_t r0 = row0 + j0;
            if (V > 1 && (fast || r0 + V <= n)) {
                u32x4 xs;
#pragma unroll
                for (int vv = 0; vv < V; ++vv) reinterpret_cast<T *>(&xs)[vv] = xv[vv];
                __builtin_amdgcn_raw_buffer_store_b128(xs, rd_rsrc, vo, j0 * rowb, cpol<NT>());
            } else {
#pragma unroll
                for (int vv = 0; vv < V; ++vv)
                    if (fast || r0 + vv < n)
                        bstore<T, NT>(xv[vv], rd_rsrc, vo, (j0 + vv) * rowb);
            }
        } else {
            const int64_t r0 = row0 + j0;
            if (V > 1 && r0 + V <= n) {
                using VT = typename Vec16<T>::type;
                VT xs;
#pragma unroll
                for (int vv = 0; vv < V; ++vv) reinterpret_cast<T *>(&xs)[vv] = xv[vv];
                *reinterpret_cast<VT *>(D + base + r0) = xs;
            } else {
#pragma unroll
                for (int vv = 0; vv < V; ++vv) {
                    const int64_t r = r0 + vv;
                    if (r < n) D[base + r * rs] = xv[vv];
                }
            }
        }
    }
}

// Naive batched Thomas: one lane per system, the recurrence of
// Tridiagonal_CUDA/tridiagonal_basic.cu:4-30.  Writes c' into c and x into d
// (the reference's in-place semantics).  Coalesced in the interleaved layout,
// lane-strided (one cache line per lane and row) in the contiguous layout.
template <typename T>
__global__ void __launch_bounds__(64) thomas_kernel(const T *__restrict__ A,
                                                    const T *__restrict__ B, T *__restrict__ C,
                                                    T *__restrict__ D, int64_t n, int64_t batch,
                                                    int64_t sstride, int64_t rs) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= batch) return;
    const int64_t base = s * sstride;
    T rb = rcp(B[base]);
    T cc = (n > 1) ? C[base] * rb : T(0);
    T dd = D[base] * rb;
    C[base] = cc;
    D[base] = dd;
    for (int64_t i = 1; i < n; ++i) {
        const int64_t o = base + i * rs;
        T a = A[o];
        T k = rcp(B[o] - cc * a);
        cc = (i < n - 1) ? C[o] * k : T(0);
        dd = (D[o] - dd * a) * k;
        C[o] = cc;
        D[o] = dd;
    }
    for (int64_t i = n - 2; i >= 0; --i) {
        const int64_t o = base + i * rs;
        dd = D[o] - C[o] * dd;
        D[o] = dd;
    }
}

}  // namespace tri
