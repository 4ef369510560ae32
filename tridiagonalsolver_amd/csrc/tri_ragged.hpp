// Ragged batch in one persistent launch.
//
// Systems of uneven size n_s (contiguous per system, concatenated; system s at
// rows [offsets[s], offsets[s+1]) of every array) are grouped on the host into
// size buckets.  Each bucket has a tile shape of the on-chip partition solver
// that fits a 512-lane workgroup (S systems per workgroup, NWS waves = 64*NWS
// lanes per system, M rows per lane) and an interface solver (Wang-Mou, PCR,
// PCR + Thomas, CR).  A work item is one workgroup-load: up to S systems of one
// bucket.  Items are ordered largest first; persistent workgroups take them
// from an atomic counter until the list is exhausted, and every workgroup
// reaches the exit once it draws an index past the end.
#pragma once
#include "tri_host.hpp"

namespace tri {

// Shapes of the ragged kernel (512 lanes): {M, NWS, S}; capacity 64*NWS*M rows.
struct RaggedShape {
    int M, NWS, S;
};
constexpr int kRaggedShapes = 12;
constexpr RaggedShape kRaggedShapeTab[kRaggedShapes] = {
    {2, 1, 8},  {4, 1, 8},  {8, 1, 8},  {8, 2, 4},  {16, 1, 8}, {8, 4, 2},
    {16, 2, 4}, {8, 8, 1},  {16, 4, 2}, {16, 8, 1}, {4, 2, 4},  {4, 4, 2}};
constexpr int ragged_cap(int k) { return 64 * kRaggedShapeTab[k].NWS * kRaggedShapeTab[k].M; }
constexpr int kRaggedLanes = 512;
constexpr int kRaggedMaxN = 64 * 8 * 16;

// Work buffer: header, then items[nitems][4] = {shape, interface solver, first
// index into perm, number of systems}, then perm[n_sys] (system ids grouped by
// bucket).  Written by the host planner once; the solve only resets `counter`.
struct RaggedHeader {
    int64_t magic;
    int64_t n_sys;
    int64_t nitems;
    int32_t counter;     // work-queue head, zeroed before every solve
    int32_t precision;
};
constexpr int64_t kRaggedMagic = 0x52414747454431LL;
constexpr int64_t kRaggedHeaderBytes = 64;
constexpr int kRaggedItemInts = 4;

template <typename T> struct RaggedParams {
    const T *a, *b, *c;
    T *d;
    const int64_t *offsets;   // device, n_sys + 1
    RaggedHeader *hdr;
    const int32_t *items;     // [nitems][4]
    const int32_t *perm;      // [n_sys]
    int64_t n_sys;
    int32_t precision;
};

__device__ __forceinline__ int64_t wave_uniform(int64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// One item of shape (M, NWS, S): the lanes of wave w belong to system slot w / NWS.
// Loads go through per-wave buffer descriptors based at the system's first row,
// bounded by its last one, with 16-byte loads and stores wherever 16 bytes of
// rows remain.
template <typename T, int M, int NWS, int S>
__device__ __forceinline__ void ragged_item(const RaggedParams<T> &p, int iface, int start,
                                            int count, char *dyn_lds) {
    constexpr bool VEC = M >= Vec16<T>::N;
    using Sh = TileShape<T, M, S * NWS, S, false, VEC, false, false, true, true, false>;
    constexpr int SW = 64, CH = Sh::CH, V = Sh::V;
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
    bool valid = sl < count && (int64_t)start + sl < p.n_sys;
    int64_t off = 0, n = 0;
    if (valid) {
        const int32_t sid = p.perm[start + sl];
        if (sid >= 0 && sid < p.n_sys) {
            off = p.offsets[sid];
            n = p.offsets[sid + 1] - off;
        }
        if (n <= 0 || n > 64 * NWS * M) { valid = false; n = 0; off = 0; }
    }
    off = wave_uniform(off);
    n = wave_uniform(n);
    const int64_t row0 = (int64_t)(ws * SW + kw) * M;
    const int32_t nb = (int32_t)(n * (int64_t)sizeof(T));
    const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.a + off, nb), rb = make_rsrc(p.b + off, nb),
                                 rc = make_rsrc(p.c + off, nb), rd = make_rsrc(p.d + off, nb);
    const int vo = (int)(row0 * (int64_t)sizeof(T));
    PredLoader<T, CH, V, true> ld{ra, rb, rc, rd, vo, (int)sizeof(T), row0, n};
    T al_r[Sh::NI];
    auto al_store = [&](int j, T v) { al_r[j] = v; };
    SweepOut<T, Sh::NI> o;
    lane_sweep<T, Sh>(o, ld, row0, n, al_store);
    const T DLr = o.bL - o.cL * o.SA, ELr = o.cL * o.PI, FLr = o.dL - o.cL * o.SD;
    const T ARr = -(o.aR * o.alp), DRr = o.bR - o.aR * o.g, FRr = o.dR - o.aR * o.del;
    T xL, xR;
    interface_solve<T, S, SW, NWS, 1, 64 * NWS, kRaggedLanes>(
        o.aL, DLr, ELr, FLr, ARr, DRr, o.cR, FRr, tid, lane, sl, kw, ws, iface, 0, dyn_lds, xL, xR);
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
#pragma unroll
    for (int j0 = 0; j0 < M; j0 += V) {
        if (V > 1 && row0 + j0 + V <= n) {
            u32x4 t;
#pragma unroll
            for (int v = 0; v < V; ++v) reinterpret_cast<T *>(&t)[v] = xs[j0 + v];
            __builtin_amdgcn_raw_buffer_store_b128(t, rd, vo, j0 * (int)sizeof(T), cpol<true>());
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (row0 + j0 + v < n) bstore<T, true>(xs[j0 + v], rd, vo, (j0 + v) * (int)sizeof(T));
        }
    }
}

// Persistent kernel.  A work buffer that does not carry this plan's magic,
// system count and precision is ignored (the launch does nothing).
template <typename T>
__global__ void __launch_bounds__(kRaggedLanes) ragged_kernel(RaggedParams<T> p) {
    extern __shared__ __attribute__((aligned(16))) char tri_dyn_lds[];
    __shared__ int s_item;
    const RaggedHeader *h = p.hdr;
    if (h->magic != kRaggedMagic || h->n_sys != p.n_sys || h->precision != p.precision) return;
    const int64_t nitems = h->nitems;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&p.hdr->counter, 1);
        __syncthreads();
        // the item index and its fields are workgroup-uniform: keep them in SGPRs
        // so that the shape switch is a scalar branch
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it < 0 || it >= nitems) break;   // every workgroup leaves once the list is empty
        const int32_t *w = p.items + (int64_t)kRaggedItemInts * it;
        const int shape = __builtin_amdgcn_readfirstlane(w[0]);
        const int iface = __builtin_amdgcn_readfirstlane(w[1]);
        const int start = __builtin_amdgcn_readfirstlane(w[2]);
        const int count = __builtin_amdgcn_readfirstlane(w[3]);
        switch (shape) {
#define TRI_RAGGED_CASE(K)                                                                   \
    case K:                                                                                \
        ragged_item<T, kRaggedShapeTab[K].M, kRaggedShapeTab[K].NWS, kRaggedShapeTab[K].S>( \
            p, iface, start, count, tri_dyn_lds);                                          \
        break;
            TRI_RAGGED_CASE(0) TRI_RAGGED_CASE(1) TRI_RAGGED_CASE(2) TRI_RAGGED_CASE(3)
            TRI_RAGGED_CASE(4) TRI_RAGGED_CASE(5) TRI_RAGGED_CASE(6) TRI_RAGGED_CASE(7)
            TRI_RAGGED_CASE(8) TRI_RAGGED_CASE(9) TRI_RAGGED_CASE(10) TRI_RAGGED_CASE(11)
#undef TRI_RAGGED_CASE
        default: break;
        }
        __syncthreads();                     // LDS of the item is free for the next one
    }
}

// Dynamic LDS of the persistent kernel: room for the LDS interface solvers.
template <typename T> constexpr int ragged_lds_bytes() { return iface_lds_bytes<T>(kRaggedLanes); }

hipError_t launch_ragged_f32(const RaggedParams<float> &p, int grid, hipStream_t st);
hipError_t launch_ragged_f64(const RaggedParams<double> &p, int grid, hipStream_t st);
int ragged_grid_f32();
int ragged_grid_f64();

}  // namespace tri
