// Adjoint (backward) of the batched solve, and the fused forward + adjoint.
//
// For x = A^{-1} d and an upstream gradient g = dL/dx:
//   lambda = A^{-T} g,  dL/da_i = -lambda_i x_{i-1},  dL/db_i = -lambda_i x_i,
//   dL/dc_i = -lambda_i x_{i+1},  dL/dd_i = lambda_i   (a_0, c_{n-1} are not in A).
// A^T is tridiagonal with sub-diagonal c_{i-1}, diagonal b_i, super-diagonal
// a_{i+1}; it is solved with the same on-chip partition solver as the forward
// problem (lane sweep, interface solve, substitution), fed with the shifted
// coefficients.
//   MODE kAdjoint : reads a, b, c, x, g; writes the four gradients (9 words/row).
//   MODE kFused   : reads a, b, c, d, g; solves A x = d, keeps a, b, c in VGPRs,
//                   solves A^T lambda = g and writes x and the gradients
//                   (10 words/row instead of 14 for a forward and a backward pass).
#pragma once
#include "tri_host.hpp"

namespace tri {

enum AdjMode : int { kAdjoint = 0, kFused = 1 };

template <typename T> struct AdjParams {
    const T *a, *b, *c;
    const T *d;   // kFused: right-hand side; kAdjoint: unused
    const T *x;   // kAdjoint: forward solution; kFused: unused
    const T *g;   // upstream gradient dL/dx
    T *xo;        // kFused: x output
    T *ga, *gb, *gc, *gd;
    int64_t n, batch, sstride, rstride;
    int32_t buf_bytes;
    int32_t iface, xchg_lds;
};

template <typename T>
__device__ __forceinline__ T rowload(__amdgpu_buffer_rsrc_t r, int vo, int64_t row, int64_t n) {
    return (row >= 0 && row < n) ? bload<T>(r, vo, 0) : T(0);
}

// Contiguous layout: a wave's rows [wrow, wrow + 64*M) of one system are read as
// whole 16-byte chunks (consecutive lanes, consecutive chunks; nontemporal) and
// handed to the lanes through the wave's LDS block `xb`.  Returns the lane's M
// rows and the values one row before (prev) and one row after (next) its
// segment, which come from the neighbouring lanes' rows or, at the block edges,
// from one extra load.  Rows at or past n read as zero.
template <typename T, int M>
__device__ __forceinline__ void adj_wave_rows(__amdgpu_buffer_rsrc_t r, int64_t wrow, int64_t n,
                                              int lane, u32x4 *xb, T (&rows)[M], T &prev,
                                              T &next, bool want_prev, bool want_next) {
    constexpr int VV = 16 / (int)sizeof(T);
    constexpr int NQ = M / VV;
    u32x4 g[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        g[q] = tail_chunk_load<T, true>(r, wrow + (int64_t)(q * 64 + lane) * VV, n);
    wave_lds_sync();                                  // the block is free
#pragma unroll
    for (int q = 0; q < NQ; ++q) xb[q * 64 + lane] = g[q];
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const u32x4 t = xb[lane * NQ + i];
#pragma unroll
        for (int v = 0; v < VV; ++v) rows[i * VV + v] = reinterpret_cast<const T *>(&t)[v];
    }
    const T *tb = reinterpret_cast<const T *>(xb);
    prev = 0;
    next = 0;
    if (want_prev) {
        prev = (lane > 0) ? tb[lane * M - 1] : T(0);
        const int64_t rp = wrow - 1;
        if (lane == 0 && rp >= 0 && rp < n) prev = bload<T, true>(r, (int)(rp * (int64_t)sizeof(T)), 0);
    }
    if (want_next) {
        next = (lane < 63) ? tb[lane * M + M] : T(0);
        const int64_t rn = wrow + 64 * M;
        if (lane == 63 && rn < n) next = bload<T, true>(r, (int)(rn * (int64_t)sizeof(T)), 0);
    }
}

// Writes the lane's M values back as whole 16-byte chunks of the wave block.
template <typename T, int M>
__device__ __forceinline__ void adj_wave_store(__amdgpu_buffer_rsrc_t r, int64_t wrow, int64_t n,
                                               int lane, u32x4 *xb, const T (&vals)[M]) {
    constexpr int VV = 16 / (int)sizeof(T);
    constexpr int NQ = M / VV;
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        u32x4 t;
#pragma unroll
        for (int v = 0; v < VV; ++v) reinterpret_cast<T *>(&t)[v] = vals[i * VV + v];
        xb[lane * NQ + i] = t;
    }
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        tail_chunk_store<T, true>(xb[q * 64 + lane], r, wrow + (int64_t)(q * 64 + lane) * VV, n);
}

template <typename T, int M, int NW, int S, bool IL, int MODE>
__global__ void __launch_bounds__(64 * NW) tile_adj_kernel(AdjParams<T> prm) {
    using Sh = TileShape<T, M, NW, S, IL, false, false, false, true, false, false>;
    constexpr int SW = Sh::SW, NWS = Sh::NWS, LS = Sh::LS, CH = Sh::CH;
    constexpr int LANES = 64 * NW;
    static_assert(IL || M * (int)sizeof(T) >= 16, "contiguous adjoint tiles move 16-byte chunks");
    extern __shared__ __attribute__((aligned(16))) char tri_dyn_lds[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int sl = IL ? (lane % S) : (wave / NWS);
    const int kw = IL ? (lane / S) : lane;
    const int ws = IL ? wave : (wave % NWS);
    const int64_t n = prm.n;
    int tile = blockIdx.x;
    if (IL) tile = xcd_group(blockIdx.x, gridDim.x, 128 / (S * (int)sizeof(T)));
    const int64_t sys = (int64_t)tile * S + sl;
    const bool valid = sys < prm.batch;
    const int64_t sysc = valid ? sys : prm.batch - 1;
    const int seg = ws * SW + kw;
    const int64_t row0 = (int64_t)seg * M;

    // raw rows j = 0..M-1 plus the neighbours needed by A^T at the segment edges
    T ra[M], rb[M], rc[M], rg[M], rx[M];
    T c_before, a_after, x_before = 0, x_after = 0;
    __shared__ u32x4 xbuf[IL ? 1 : NW][IL ? 1 : 64 * M * (int)sizeof(T) / 16];
    (void)xbuf;
    // contiguous: per-wave descriptors on the system's own rows
    const int64_t sbase = IL ? 0 : wave_uniform(sysc * n);
    const int64_t nn = valid ? n : 0;
    const int64_t wrow = (int64_t)ws * 64 * M;
    auto RW = [&](const T *p) { return make_rsrc(p + sbase, (int32_t)(n * (int64_t)sizeof(T))); };
    if constexpr (IL) {
        const int64_t rs = prm.rstride;
        const int nb = prm.buf_bytes;
        const int rowb = (int)(rs * (int64_t)sizeof(T));
        const int vo0 = (int)((sysc + row0 * rs) * (int64_t)sizeof(T));
        auto RS = [&](const T *p) { return make_rsrc(p, nb); };
        const __amdgpu_buffer_rsrc_t rA = RS(prm.a), rB = RS(prm.b), rC = RS(prm.c);
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const int64_t r = row0 + j;
            const int vo = vo0 + j * rowb;
            rb[j] = rowload<T>(rB, vo, r, n);
            rc[j] = rowload<T>(rC, vo, r, n);
            ra[j] = rowload<T>(rA, vo, r, n);
            rg[j] = rowload<T>(RS(prm.g), vo, r, n);
            if constexpr (MODE == kFused) rx[j] = rowload<T>(RS(prm.d), vo, r, n);
            else rx[j] = rowload<T>(RS(prm.x), vo, r, n);
        }
        c_before = rowload<T>(rC, vo0 - rowb, row0 - 1, n);      // c_{row0-1}
        a_after = rowload<T>(rA, vo0 + M * rowb, row0 + M, n);   // a_{row0+M}
        if constexpr (MODE == kAdjoint) {
            x_before = rowload<T>(RS(prm.x), vo0 - rowb, row0 - 1, n);
            x_after = rowload<T>(RS(prm.x), vo0 + M * rowb, row0 + M, n);
        }
    } else {
        u32x4 *xb = xbuf[wave];
        T dummy0, dummy1;
        adj_wave_rows<T, M>(RW(prm.b), wrow, nn, lane, xb, rb, dummy0, dummy1, false, false);
        adj_wave_rows<T, M>(RW(prm.c), wrow, nn, lane, xb, rc, c_before, dummy1, true, false);
        adj_wave_rows<T, M>(RW(prm.a), wrow, nn, lane, xb, ra, dummy0, a_after, false, true);
        adj_wave_rows<T, M>(RW(prm.g), wrow, nn, lane, xb, rg, dummy0, dummy1, false, false);
        if constexpr (MODE == kFused)
            adj_wave_rows<T, M>(RW(prm.d), wrow, nn, lane, xb, rx, dummy0, dummy1, false, false);
        else
            adj_wave_rows<T, M>(RW(prm.x), wrow, nn, lane, xb, rx, x_before, x_after, true, true);
    }
#pragma unroll
    for (int j = 0; j < M; ++j)
        if (row0 + j >= n) rb[j] = 1;                    // identity rows past the end

    T al_r[Sh::NI];
    auto al_store = [&](int j, T v) { al_r[j] = v; };

    // ------------------------------------------------------------ forward (fused)
    if constexpr (MODE == kFused) {
        T rows[4][M];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            rows[0][j] = ra[j]; rows[1][j] = rb[j]; rows[2][j] = rc[j]; rows[3][j] = rx[j];
        }
        SweepOut<T, Sh::NI> o;
        RegLoader<T, CH, M> ld{rows};
        lane_sweep<T, Sh>(o, ld, row0, n, al_store);
        const T DLr = o.bL - o.cL * o.SA, ELr = o.cL * o.PI, FLr = o.dL - o.cL * o.SD;
        const T ARr = -(o.aR * o.alp), DRr = o.bR - o.aR * o.g, FRr = o.dR - o.aR * o.del;
        T xL, xR;
        interface_solve<T, S, SW, NWS, LS, Sh::P, LANES>(o.aL, DLr, ELr, FLr, ARr, DRr, o.cR, FRr,
                                                         tid, lane, sl, kw, ws, prm.iface,
                                                         prm.xchg_lds, tri_dyn_lds, xL, xR);
        rx[0] = xL;
        rx[M - 1] = xR;
        T xn = xR;
#pragma unroll
        for (int j = M - 3; j >= 0; --j) {
            xn = (o.de[j] - al_r[j] * xL) - o.ga[j] * xn;
            rx[j + 1] = xn;
        }
        // x at the rows just outside the segment, from the neighbouring segments
        __shared__ T sh_x[S][2 * Sh::P];
        sh_x[sl][2 * seg] = xL;
        sh_x[sl][2 * seg + 1] = xR;
        __syncthreads();
        x_before = (seg > 0) ? sh_x[sl][2 * seg - 1] : T(0);
        x_after = (seg + 1 < Sh::P) ? sh_x[sl][2 * seg + 2] : T(0);
        __syncthreads();   // interface LDS is reused by the adjoint solve
    }

    // ------------------------------------------------------------ adjoint solve
    T lam[M];
    {
        T rows[4][M];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            rows[0][j] = (j == 0) ? c_before : rc[j - 1];      // sub of A^T: c_{i-1}
            rows[1][j] = rb[j];
            rows[2][j] = (j == M - 1) ? a_after : ra[j + 1];   // super of A^T: a_{i+1}
            rows[3][j] = rg[j];
            if (row0 + j >= n) { rows[0][j] = 0; rows[2][j] = 0; rows[3][j] = 0; }
        }
        SweepOut<T, Sh::NI> o;
        RegLoader<T, CH, M> ld{rows};
        lane_sweep<T, Sh>(o, ld, row0, n, al_store);
        const T DLr = o.bL - o.cL * o.SA, ELr = o.cL * o.PI, FLr = o.dL - o.cL * o.SD;
        const T ARr = -(o.aR * o.alp), DRr = o.bR - o.aR * o.g, FRr = o.dR - o.aR * o.del;
        T lL, lR;
        interface_solve<T, S, SW, NWS, LS, Sh::P, LANES>(o.aL, DLr, ELr, FLr, ARr, DRr, o.cR, FRr,
                                                         tid, lane, sl, kw, ws, prm.iface,
                                                         prm.xchg_lds, tri_dyn_lds, lL, lR);
        lam[0] = lL;
        lam[M - 1] = lR;
        T ln = lR;
#pragma unroll
        for (int j = M - 3; j >= 0; --j) {
            ln = (o.de[j] - al_r[j] * lL) - o.ga[j] * ln;
            lam[j + 1] = ln;
        }
    }
    if (!valid) return;

    // ------------------------------------------------------------ gradients
    // ga_i = -lambda_i x_{i-1}, gb_i = -lambda_i x_i, gc_i = -lambda_i x_{i+1}, gd_i = lambda_i
    auto grad = [&](int k, int j) -> T {
        const int64_t r = row0 + j;
        const T xm = (j == 0) ? x_before : rx[j - 1];
        const T xp = (j == M - 1) ? x_after : rx[j + 1];
        const T l = lam[j];
        if (k == 0) return (r > 0) ? -(l * xm) : T(0);
        if (k == 1) return -(l * rx[j]);
        if (k == 2) return (r < n - 1) ? -(l * xp) : T(0);
        if (k == 3) return l;
        return rx[j];
    };
    T *outs[5] = {prm.ga, prm.gb, prm.gc, prm.gd, prm.xo};
    constexpr int NOUT = (MODE == kFused) ? 5 : 4;
    if constexpr (IL) {
        const int64_t rs = prm.rstride;
        const int rowb = (int)(rs * (int64_t)sizeof(T));
        const int vo0 = (int)((sysc + row0 * rs) * (int64_t)sizeof(T));
#pragma unroll
        for (int k = 0; k < NOUT; ++k) {
            const __amdgpu_buffer_rsrc_t w = make_rsrc(outs[k], prm.buf_bytes);
#pragma unroll
            for (int j = 0; j < M; ++j)
                if (row0 + j < n) bstore<T>(grad(k, j), w, vo0 + j * rowb, 0);
        }
    } else {
        if (wrow >= n) return;                             // wave-uniform
#pragma unroll
        for (int k = 0; k < NOUT; ++k) {
            T v[M];
#pragma unroll
            for (int j = 0; j < M; ++j) v[j] = grad(k, j);
            adj_wave_store<T, M>(RW(outs[k]), wrow, n, lane, xbuf[wave], v);
        }
    }
}

template <typename T> struct AdjEntry {
    int M, NW, S;
    bool IL;
    int mode;
    int cap;
    hipError_t (*launch)(const AdjParams<T> &, hipStream_t);
};

template <typename T, int M, int NW, int S, bool IL, int MODE>
hipError_t launch_adj(const AdjParams<T> &p, hipStream_t st) {
    const int64_t tiles = (p.batch + S - 1) / S;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffff) return hipErrorInvalidValue;
    auto kern = tile_adj_kernel<T, M, NW, S, IL, MODE>;
    size_t dyn = 0;
    if (iface_family(p.iface) >= kIfacePCR) dyn = (size_t)iface_lds_bytes<T>(64 * NW);
    else if (p.xchg_lds) dyn = (size_t)6 * 64 * NW * sizeof(T);
    if (dyn) {
        static size_t granted = 0;
        if (dyn > granted) {
            hipFuncAttributes fa;
            hipError_t e = hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(kern));
            if (e != hipSuccess) return e;
            if (fa.sharedSizeBytes + dyn > (size_t)max_lds_per_block()) return hipErrorInvalidConfiguration;
            e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
            if (e != hipSuccess) return e;
            granted = dyn;
        }
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)tiles), dim3(64 * NW), dyn, st, p);
    return hipGetLastError();
}

template <typename T> struct AdjTable {
    const AdjEntry<T> *entries;
    int count;
};
AdjTable<float> adj_table_f32();
AdjTable<double> adj_table_f64();
template <typename T> AdjTable<T> adj_table();
template <> inline AdjTable<float> adj_table<float>() { return adj_table_f32(); }
template <> inline AdjTable<double> adj_table<double>() { return adj_table_f64(); }

}  // namespace tri
