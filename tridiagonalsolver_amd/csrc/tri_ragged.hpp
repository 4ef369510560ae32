// Ragged batch in one persistent launch.
//
// Systems of uneven size n_s (contiguous per system, concatenated; system s at
// rows [offsets[s], offsets[s+1]) of every array) are grouped on the host into
// size buckets.  Each bucket has a tile shape of the on-chip partition solver
// that fits a 512-lane workgroup (S systems per workgroup, NWS waves = 64*NWS
// lanes per system, M rows per lane) and an interface solver (Wang-Mou, PCR,
// PCR + Thomas, CR).  A work item is one workgroup-load: up to S systems of one
// bucket, described by a 128-byte record (shape, interface solver and the first
// row and size of each system).  Items are ordered largest first; persistent
// workgroups take them from an atomic counter.  The index and record of the next
// item are fetched while the current one is solved, so a workgroup starts the
// coefficient loads of an item without a dependent chain of global reads.  Every
// workgroup leaves the loop once it draws an index past the end of the list.
#pragma once
#include "tri_host.hpp"
#include "tri_ragged_plan.hpp"

namespace tri {

template <typename T> struct RaggedParams {
    const T *a, *b, *c;
    T *d;
    RaggedHeader *hdr;
    const int32_t *recs;      // [nitems][32]
    int64_t n_sys;
    int32_t precision;
};

// One item of shape (M, NWS, S): the lanes of wave w belong to system slot w / NWS
// and wave w % NWS of that system owns rows [64*M*(w % NWS), 64*M*(w % NWS + 1)).
// Segments of at least 16 bytes per lane are read as whole 16-byte chunks of
// the wave's row block (consecutive lanes, consecutive chunks) and handed to
// the lanes through LDS; shorter segments are read row by row.
template <typename T, int M, int NWS, int S>
__device__ __forceinline__ void ragged_item(const RaggedParams<T> &p, const int32_t *rec,
                                            int iface, int count, int64_t rows, u32x4 *xbuf,
                                            char *iface_lds) {
    constexpr int VV = 16 / (int)sizeof(T);
    constexpr bool XPL = M * (int)sizeof(T) >= 16;          // whole-chunk loads
    constexpr int NCH = XPL ? 64 * M / VV : 1;               // 16-byte chunks per wave block
    constexpr int NQ = XPL ? NCH / 64 : 1;                   // chunks per lane
    constexpr int PL = XPL ? M / VV : 1;                     // chunks per lane segment
    static_assert(!XPL || M % VV == 0, "whole-chunk loads need whole chunks per lane");
    using Sh = TileShape<T, M, S * NWS, S, false, false, false, false, true, true, XPL>;
    constexpr int SW = 64, CH = Sh::CH;
    // Opaque copy of the lane id: without it the lane-dependent values of every
    // shape are hoisted out of the persistent loop and stay live in VGPRs across
    // all items of every shape.
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int sl = wave / NWS, ws = wave % NWS, kw = lane;
    // slots past `count` are idle systems (identity rows) that still take part
    // in the workgroup-wide interface solve
    const bool valid = sl < count;
    int64_t off = 0, n = 0;
    if (valid) {
        off = (int64_t)(((uint64_t)(uint32_t)rec[5 + 3 * sl] << 32) | (uint32_t)rec[4 + 3 * sl]);
        n = rec[6 + 3 * sl];
    }
    off = wave_uniform(off);
    n = wave_uniform(n);
    const int64_t row0 = (int64_t)(ws * SW + kw) * M;
    const int64_t wrow = (int64_t)ws * 64 * M;               // first row of the wave block
    const int32_t nb = (int32_t)(n * (int64_t)sizeof(T));
    const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.a + off, nb), rb = make_rsrc(p.b + off, nb),
                                 rc = make_rsrc(p.c + off, nb), rd = make_rsrc(p.d + off, nb);
    T al_r[Sh::NI];
    auto al_store = [&](int j, T v) { al_r[j] = v; };
    SweepOut<T, Sh::NI> o;
    if constexpr (XPL) {
        // Whole chunks are read without a bounds branch, from resources that reach to
        // the last row of the arrays: a branch per chunk ends in a copy that waits for
        // the load in flight, which serialised the loads of every padded system.
        // Chunks past the system read the system's first chunk again (a cache hit; its
        // rows become identity rows below).  Only a chunk that holds rows of the system
        // and runs past the last row of the arrays takes element loads, once all others
        // are in flight.
        // (idle slots: an empty range, whose loads return zeros without reading memory)
        const int64_t rem = wave_uniform(!valid ? 0
                                         : (rows - off < (int64_t)kRaggedMaxN) ? rows - off
                                                                                : kRaggedMaxN);
        const int32_t nr = (int32_t)(rem * (int64_t)sizeof(T));
        const __amdgpu_buffer_rsrc_t rr[4] = {make_rsrc(p.a + off, nr), make_rsrc(p.b + off, nr),
                                              make_rsrc(p.c + off, nr), make_rsrc(p.d + off, nr)};
        u32x4 g[4][NQ];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int64_t e = wrow + (int64_t)(q * 64 + lane) * VV;
                const int vo = (e < n && e + VV <= rem) ? (int)(e * (int64_t)sizeof(T)) : 0;
                g[k][q] = __builtin_amdgcn_raw_buffer_load_b128(rr[k], vo, 0, cpol<true>());
            }
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int64_t e = wrow + (int64_t)(q * 64 + lane) * VV;
                if (e < n && e + VV > rem) g[k][q] = tail_chunk_load<T, true>(rr[k], e, n);
            }
        T rows[4][M];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k > 0) wave_lds_sync();          // previous array fully read by all lanes
#pragma unroll
            for (int q = 0; q < NQ; ++q) xbuf[q * 64 + lane] = g[k][q];
            wave_lds_sync();                     // block written before lanes read their rows
#pragma unroll
            for (int i = 0; i < PL; ++i) {
                const u32x4 t = xbuf[lane * PL + i];
#pragma unroll
                for (int v = 0; v < VV; ++v) rows[k][i * VV + v] = reinterpret_cast<const T *>(&t)[v];
            }
        }
#pragma unroll
        for (int j = 0; j < M; ++j)
            if (row0 + j >= n) { rows[0][j] = 0; rows[1][j] = 1; rows[2][j] = 0; rows[3][j] = 0; }
        RegLoader<T, CH, M> ld{rows};
        lane_sweep<T, Sh>(o, ld, row0, n, al_store);
    } else {
        PredLoader<T, CH, 1, true> ld{ra, rb, rc, rd, (int)(row0 * (int64_t)sizeof(T)),
                                      (int)sizeof(T), row0, n};
        lane_sweep<T, Sh>(o, ld, row0, n, al_store);
    }
    const T DLr = o.bL - o.cL * o.SA, ELr = o.cL * o.PI, FLr = o.dL - o.cL * o.SD;
    const T ARr = -(o.aR * o.alp), DRr = o.bR - o.aR * o.g, FRr = o.dR - o.aR * o.del;
    T xL, xR;
    interface_solve<T, S, SW, NWS, 1, 64 * NWS, kRaggedLanes>(
        o.aL, DLr, ELr, FLr, ARr, DRr, o.cR, FRr, tid, lane, sl, kw, ws, iface, 0, iface_lds, xL, xR);
    if (!valid || wrow >= n) return;                         // wave-uniform
    T xs[M];
    xs[0] = xL;
    xs[M - 1] = xR;
    T xn = xR;
#pragma unroll
    for (int j = M - 3; j >= 0; --j) {
        xn = (o.de[j] - al_r[j] * xL) - o.ga[j] * xn;
        xs[j + 1] = xn;
    }
    if constexpr (XPL) {
        wave_lds_sync();
#pragma unroll
        for (int i = 0; i < PL; ++i) {
            u32x4 t;
#pragma unroll
            for (int v = 0; v < VV; ++v) reinterpret_cast<T *>(&t)[v] = xs[i * VV + v];
            xbuf[lane * PL + i] = t;
        }
        wave_lds_sync();
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            tail_chunk_store<T, true>(xbuf[q * 64 + lane], rd, wrow + (int64_t)(q * 64 + lane) * VV, n);
    } else {
        const int vo = (int)(row0 * (int64_t)sizeof(T));
#pragma unroll
        for (int j = 0; j < M; ++j)
            if (row0 + j < n) bstore<T, true>(xs[j], rd, vo, j * (int)sizeof(T));
    }
}

// Rows j0 .. j0+CH-1 of the lane's own system through 64-bit addresses (the
// systems of one wave differ, so no wave-uniform buffer descriptor covers them):
// whole 16-byte vectors where they lie inside the system (4-byte aligned, as the
// systems start anywhere), single elements at its end, identity rows past it.
template <typename T, int CH> struct LaneRowLoader {
    const T *a, *b, *c, *d;      // row 0 of the lane's system
    int64_t row0, n;
    __device__ __forceinline__ void load(T (&q)[4][CH], int j0) const {
        constexpr int V = 16 / (int)sizeof(T);
        static_assert(CH % V == 0, "rows move in 16-byte vectors");
        typedef unsigned int u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));
        const T *src[4] = {a, b, c, d};
#pragma unroll
        for (int v0 = 0; v0 < CH; v0 += V) {
            const int64_t r = row0 + j0 + v0;
            if (r + V <= n) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const u32x4 t = *reinterpret_cast<const u32x4a4 *>(src[k] + r);
#pragma unroll
                    for (int v = 0; v < V; ++v) q[k][v0 + v] = reinterpret_cast<const T *>(&t)[v];
                }
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    if (r + v < n) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) q[k][v0 + v] = src[k][r + v];
                    } else {
                        q[0][v0 + v] = 0; q[1][v0 + v] = 1; q[2][v0 + v] = 0; q[3][v0 + v] = 0;
                    }
                }
            }
        }
    }
};

// One item of a sub-wave shape (M rows per lane, S systems of L = 512/S lanes,
// LW = 64/L systems per wave).  Lane = kw*LW + the system's place in its wave, the
// lane order of the interleaved tiles, so that the Wang-Mou merges and the LDS
// interface solvers run as they are with L segments per system and no cross-wave
// chain.  The item's systems are entries tab[0 .. count) of the system table.
// Short systems in the wave shapes leave most lanes holding padding rows and pay
// a workgroup-wide solve per 8 systems; here a 512-lane item holds 32 or 64.
template <typename T, int M, int S>
__device__ __forceinline__ void ragged_item_sub(const RaggedParams<T> &p, const int32_t *tab,
                                                int iface, int count, char *iface_lds) {
    constexpr int L = kRaggedLanes / S;
    constexpr int LW = 64 / L;
    static_assert(L * S == kRaggedLanes && LW * L == 64, "sub-wave shape");
    using Sh = TileShape<T, M, 8, 8, false, true, false, false, true, true, false>;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wave = tid >> 6;
    const int kw = lane / LW, slw = lane % LW;
    const int sl = wave * LW + slw;
    const bool valid = sl < count;
    int64_t off = 0, n = 0;
    if (valid) {
        const int32_t *e = tab + (int64_t)kRaggedTabInts * sl;
        off = (int64_t)(((uint64_t)(uint32_t)e[1] << 32) | (uint32_t)e[0]);
        n = e[2];
    }
    const int64_t row0 = (int64_t)kw * M;
    T al_r[Sh::NI];
    auto al_store = [&](int j, T v) { al_r[j] = v; };
    SweepOut<T, Sh::NI> o;
    LaneRowLoader<T, Sh::CH> ld{p.a + off, p.b + off, p.c + off, p.d + off, row0, n};
    lane_sweep<T, Sh>(o, ld, row0, n, al_store);
    const T DLr = o.bL - o.cL * o.SA, ELr = o.cL * o.PI, FLr = o.dL - o.cL * o.SD;
    const T ARr = -(o.aR * o.alp), DRr = o.bR - o.aR * o.g, FRr = o.dR - o.aR * o.del;
    T xL, xR;
    interface_solve<T, S, L, 1, LW, L, kRaggedLanes>(
        o.aL, DLr, ELr, FLr, ARr, DRr, o.cR, FRr, tid, lane, sl, kw, 0, iface, 0, iface_lds, xL, xR);
    if (!valid || row0 >= n) return;
    T xs[M];
    xs[0] = xL;
    xs[M - 1] = xR;
    T xn = xR;
#pragma unroll
    for (int j = M - 3; j >= 0; --j) {
        xn = (o.de[j] - al_r[j] * xL) - o.ga[j] * xn;
        xs[j + 1] = xn;
    }
    constexpr int V = 16 / (int)sizeof(T);
    typedef unsigned int u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));
    T *x = p.d + off;
#pragma unroll
    for (int v0 = 0; v0 < M; v0 += V) {
        const int64_t r = row0 + v0;
        if (r + V <= n) {
            u32x4 t;
#pragma unroll
            for (int v = 0; v < V; ++v) reinterpret_cast<T *>(&t)[v] = xs[v0 + v];
            *reinterpret_cast<u32x4a4 *>(x + r) = t;
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (r + v < n) x[r + v] = xs[v0 + v];
        }
    }
}

// Dynamic LDS of the persistent kernel: the per-wave transposition blocks, then
// room for the LDS interface solvers.
template <typename T> constexpr int ragged_xbuf_bytes() {
    return kRaggedWaves * 64 * kRaggedMaxM * (int)sizeof(T);
}
template <typename T> constexpr int ragged_lds_bytes() {
    return ragged_xbuf_bytes<T>() + iface_lds_bytes<T>(kRaggedLanes);
}

// Persistent kernel.  A work buffer that does not carry this plan's magic,
// system count and precision is ignored (the launch does nothing).
template <typename T>
__global__ void __launch_bounds__(kRaggedLanes) ragged_kernel(RaggedParams<T> p) {
    extern __shared__ __attribute__((aligned(16))) char tri_dyn_lds[];
    __shared__ int32_t s_rec[2][kRaggedRecInts];
    __shared__ int32_t s_idx[2];
    const RaggedHeader *h = p.hdr;
    if (h->magic != kRaggedMagic || h->n_sys != p.n_sys || h->precision != p.precision) return;
    const int nitems = (int)h->nitems;
    const int64_t rows = h->rows;
    const int32_t *tab = p.recs + (int64_t)kRaggedRecInts * nitems;   // system table
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    u32x4 *xbuf = reinterpret_cast<u32x4 *>(tri_dyn_lds) +
                  (size_t)wave * (64 * kRaggedMaxM * sizeof(T) / 16);
    char *iface_lds = tri_dyn_lds + ragged_xbuf_bytes<T>();
    // Wave 0 keeps the queue two items ahead: the record of the next item and
    // the index of the one after it are in flight while an item is solved.
    int nxt = 0;
    if (wave == 0) {
        int i0 = 0, i1 = 0;
        if (lane == 0) {
            i0 = atomicAdd(&p.hdr->counter, 1);
            i1 = atomicAdd(&p.hdr->counter, 1);
        }
        i0 = __builtin_amdgcn_readfirstlane(i0);
        i1 = __builtin_amdgcn_readfirstlane(i1);
        if (lane < kRaggedRecInts)
            s_rec[0][lane] = (i0 < nitems) ? p.recs[(int64_t)i0 * kRaggedRecInts + lane] : 0;
        if (lane == 0) s_idx[0] = i0;
        nxt = i1;
    }
    __syncthreads();
    for (int k = 0;; ++k) {
        const int buf = k & 1;
        const int it = __builtin_amdgcn_readfirstlane(s_idx[buf]);
        if (it < 0 || it >= nitems) break;   // every workgroup leaves once the list is empty
        // wave 0: start fetching the next record and the index after it
        int32_t rnext = 0;
        int nn = 0;
        if (wave == 0) {
            if (nxt < nitems && lane < kRaggedRecInts)
                rnext = p.recs[(int64_t)nxt * kRaggedRecInts + lane];
            if (lane == 0 && nxt < nitems) nn = atomicAdd(&p.hdr->counter, 1);
        }
        const int32_t *rec = s_rec[buf];
        const int shape = __builtin_amdgcn_readfirstlane(rec[0]);
        const int iface = __builtin_amdgcn_readfirstlane(rec[1]);
        const int count = __builtin_amdgcn_readfirstlane(rec[2]);
        const int first = __builtin_amdgcn_readfirstlane(rec[3]);
        switch (shape) {
#define TRI_RAGGED_CASE(K)                                                                   \
    case K:                                                                                \
        if constexpr (kRaggedShapeTab[K].NWS == 0)                                         \
            ragged_item_sub<T, kRaggedShapeTab[K].M, kRaggedShapeTab[K].S>(                \
                p, tab + (int64_t)kRaggedTabInts * first, iface, count, iface_lds);        \
        else                                                                               \
            ragged_item<T, kRaggedShapeTab[K].M, kRaggedShapeTab[K].NWS, kRaggedShapeTab[K].S>( \
                p, rec, iface, count, rows, xbuf, iface_lds);                              \
        break;
            TRI_RAGGED_CASE(0) TRI_RAGGED_CASE(1) TRI_RAGGED_CASE(2) TRI_RAGGED_CASE(3)
            TRI_RAGGED_CASE(4) TRI_RAGGED_CASE(5) TRI_RAGGED_CASE(6) TRI_RAGGED_CASE(7)
            TRI_RAGGED_CASE(8) TRI_RAGGED_CASE(9) TRI_RAGGED_CASE(10) TRI_RAGGED_CASE(11)
            TRI_RAGGED_CASE(12) TRI_RAGGED_CASE(13) TRI_RAGGED_CASE(14) TRI_RAGGED_CASE(15)
            TRI_RAGGED_CASE(16) TRI_RAGGED_CASE(17) TRI_RAGGED_CASE(18) TRI_RAGGED_CASE(19)
            TRI_RAGGED_CASE(20) TRI_RAGGED_CASE(21) TRI_RAGGED_CASE(22) TRI_RAGGED_CASE(23)
            TRI_RAGGED_CASE(24)
#undef TRI_RAGGED_CASE
        default: break;
        }
        static_assert(kRaggedShapes == 25, "one case per shape");
        if (wave == 0) {
            if (lane < kRaggedRecInts) s_rec[buf ^ 1][lane] = rnext;
            if (lane == 0) s_idx[buf ^ 1] = nxt;
            nn = __builtin_amdgcn_readfirstlane(nn);
            nxt = (nxt < nitems) ? nn : nxt;
        }
        __syncthreads();                     // next record visible; LDS of the item free
    }
    // Thread 0 made every queue increment of this workgroup, each completed (its
    // value was used) before this point.  The workgroup that arrives last therefore
    // sees every other workgroup's increments done and resets the queue for the next
    // solve, which is thus a single launch (no reset kernel or memset to replay).
    // Relaxed order suffices (the returned values order the atomics); an acq_rel
    // increment would write back and invalidate the L2 of every workgroup's XCD.
    if (tid == 0) {
        const int32_t before = __hip_atomic_fetch_add(&p.hdr->done, 1, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
        if (before == (int32_t)gridDim.x - 1) {
            __hip_atomic_store(&p.hdr->counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&p.hdr->done, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

hipError_t launch_ragged_f32(const RaggedParams<float> &p, int grid, hipStream_t st);
hipError_t launch_ragged_f64(const RaggedParams<double> &p, int grid, hipStream_t st);
int ragged_grid_f32();
int ragged_grid_f64();

}  // namespace tri
