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

template <typename T, int M, int NW, int S, bool IL, int MODE>
__global__ void __launch_bounds__(64 * NW) tile_adj_kernel(AdjParams<T> prm) {
    using Sh = TileShape<T, M, NW, S, IL, false, false, false, true, false, false>;
    constexpr int SW = Sh::SW, NWS = Sh::NWS, LS = Sh::LS, CH = Sh::CH;
    constexpr int LANES = 64 * NW;
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
    const int64_t rs = prm.rstride;
    const int64_t base = sysc * prm.sstride;
    const int nb = prm.buf_bytes;
    const int rowb = (int)(rs * (int64_t)sizeof(T));
    // lane byte offset of row (row0 + j) is vo0 + j * rowb
    const int vo0 = (int)((base + (IL ? row0 * rs : row0)) * (int64_t)sizeof(T));
    auto RS = [&](const T *p) { return make_rsrc(p, nb); };
    const __amdgpu_buffer_rsrc_t rA = RS(prm.a), rB = RS(prm.b), rC = RS(prm.c);

    // raw rows j = 0..M-1 plus the neighbours needed by A^T at the segment edges
    T ra[M], rb[M], rc[M], rg[M], rx[M];
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const int64_t r = row0 + j;
        const int vo = vo0 + j * rowb;
        rb[j] = rowload<T>(rB, vo, r, n);
        rc[j] = rowload<T>(rC, vo, r, n);
        ra[j] = rowload<T>(rA, vo, r, n);
        rg[j] = rowload<T>(RS(prm.g), vo, r, n);
        if constexpr (MODE == kFused) rx[j] = (r < n) ? rowload<T>(RS(prm.d), vo, r, n) : T(0);
        else rx[j] = rowload<T>(RS(prm.x), vo, r, n);
        if (r >= n) rb[j] = 1;                       // identity rows past the end
    }
    const T c_before = rowload<T>(rC, vo0 - rowb, row0 - 1, n);   // c_{row0-1}
    const T a_after = rowload<T>(rA, vo0 + M * rowb, row0 + M, n);  // a_{row0+M}
    T x_before = 0, x_after = 0;
    if constexpr (MODE == kAdjoint) {
        x_before = rowload<T>(RS(prm.x), vo0 - rowb, row0 - 1, n);
        x_after = rowload<T>(RS(prm.x), vo0 + M * rowb, row0 + M, n);
    }

    __shared__ T sh_al[1][1];
    (void)sh_al;
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
    const __amdgpu_buffer_rsrc_t wA = RS(prm.ga), wB = RS(prm.gb), wC = RS(prm.gc), wD = RS(prm.gd);
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const int64_t r = row0 + j;
        if (r >= n) break;
        const int vo = vo0 + j * rowb;
        const T xm = (j == 0) ? x_before : rx[j - 1];
        const T xp = (j == M - 1) ? x_after : rx[j + 1];
        const T l = lam[j];
        bstore<T>((r > 0) ? -(l * xm) : T(0), wA, vo, 0);
        bstore<T>(-(l * rx[j]), wB, vo, 0);
        bstore<T>((r < n - 1) ? -(l * xp) : T(0), wC, vo, 0);
        bstore<T>(l, wD, vo, 0);
        if constexpr (MODE == kFused) bstore<T>(rx[j], RS(prm.xo), vo, 0);
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
    if (p.iface >= kIfacePCR) dyn = (size_t)iface_lds_bytes<T>(64 * NW);
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
