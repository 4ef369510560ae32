// Systems larger than one workgroup holds: three passes over chunks of C rows.
//
//   pass 1 (chunk_kernel<.., kChunkReduce>): every workgroup sweeps one chunk of
//     one system with the on-chip partition solver (lane sweeps, Wang-Mou merges
//     inside each wave, then a fold over the waves) and writes the chunk's first
//     row F and last row E in outer form
//         x_F = rF - pF u - qF v ,   x_E = rE - pE u - qE v ,
//     u = x of the row before the chunk, v = x of the row after it (6 values).
//   pass 2 (chunk_iface_kernel): one workgroup per system merges the K chunk
//     outer forms pairwise in log2(K) Wang-Mou rounds; the merged block is the
//     whole system (u = v = 0), so every chunk's x_F and x_E is its r.  It writes
//     each chunk's boundary values u_k = x_{E_{k-1}}, v_k = x_{F_{k+1}}.
//   pass 3 (chunk_kernel<.., kChunkSolve>): every workgroup reads its chunk again,
//     folds the known boundary values into its first and last right-hand sides
//     (d_F -= a_F u_k, d_E -= c_E v_k) and solves the chunk as an independent
//     system with the tile solver.
// HBM traffic: 4 words per row in pass 1, 4 + 1 in pass 3, plus 8 words per
// chunk of scratch: 9 words per row against the 5 of an on-chip solve.
#pragma once
#include "tri_host.hpp"

namespace tri {

enum ChunkPass : int { kChunkReduce = 1, kChunkSolve = 3 };

template <typename T> struct LargeParams {
    const T *a, *b, *c;
    T *d;
    int64_t n;          // rows per system
    int64_t batch;      // systems in this launch
    int64_t sstride;    // elements between consecutive systems
    int32_t K;          // chunks per system
    int32_t iface;      // interface solver of pass 3
    T *red;             // [batch][K][6] chunk outer forms (pass 1 -> pass 2)
    T *bnd;             // [batch][K][2] boundary values u_k, v_k (pass 2 -> pass 3)
};

// Chunk k of system s (blockIdx.x = k, blockIdx.y = s) holds rows
// [k*C, min((k+1)*C, n)), C = 64*NW*M; wave w owns rows [64*M*w, 64*M*(w+1)).
template <typename T, int M, int NW, int PASS>
__global__ void __launch_bounds__(64 * NW) chunk_kernel(LargeParams<T> p) {
    using Sh = TileShape<T, M, NW, 1, false, false, false, false, true, true, true>;
    constexpr int VV = 16 / (int)sizeof(T);
    constexpr int NQ = M / VV;
    constexpr int C = 64 * NW * M;
    constexpr int LANES = 64 * NW;
    static_assert(M % VV == 0, "chunk segments move whole 16-byte chunks");
    extern __shared__ __attribute__((aligned(16))) char tri_dyn_lds[];
    __shared__ u32x4 xbuf[NW][64 * M * (int)sizeof(T) / 16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int k = blockIdx.x, K = p.K;
    const int64_t s = blockIdx.y;
    const int64_t first = (int64_t)k * C;
    const int64_t len = (p.n - first < C) ? p.n - first : C;
    const int64_t base = s * p.sstride + first;
    const int32_t nb = (int32_t)(len * (int64_t)sizeof(T));
    const __amdgpu_buffer_rsrc_t rr[4] = {make_rsrc(p.a + base, nb), make_rsrc(p.b + base, nb),
                                          make_rsrc(p.c + base, nb), make_rsrc(p.d + base, nb)};
    const int64_t wrow = (int64_t)wave * 64 * M;
    const int64_t row0 = (int64_t)(wave * 64 + lane) * M;   // chunk-relative first row
    u32x4 *xb = xbuf[wave];

    // whole 16-byte chunks of the wave block, handed to the lanes through LDS;
    // pass 1 reads with the default cache policy so that pass 3 can find the
    // chunk in the last-level cache, pass 3 streams
    u32x4 g[4][NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int a = 0; a < 4; ++a)
            g[a][q] = tail_chunk_load<T, PASS == kChunkSolve>(
                rr[a], wrow + (int64_t)(q * 64 + lane) * VV, len);
    T rows[4][M];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        if (a > 0) wave_lds_sync();
#pragma unroll
        for (int q = 0; q < NQ; ++q) xb[q * 64 + lane] = g[a][q];
        wave_lds_sync();
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const u32x4 t = xb[lane * NQ + i];
#pragma unroll
            for (int v = 0; v < VV; ++v) rows[a][i * VV + v] = reinterpret_cast<const T *>(&t)[v];
        }
    }
    // identity rows past the chunk; the system's own first and last rows are
    // uncoupled (a_0 and c_{n-1} are not part of A)
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const int64_t r = row0 + j;
        if (r >= len) { rows[0][j] = 0; rows[1][j] = 1; rows[2][j] = 0; rows[3][j] = 0; }
        if (k == 0 && r == 0) rows[0][j] = 0;
        if (k == K - 1 && r == len - 1) rows[2][j] = 0;
    }
    if constexpr (PASS == kChunkSolve) {
        // fold the chunk's neighbours into its end rows: an independent system
        const T u = p.bnd[(s * K + k) * 2], v = p.bnd[(s * K + k) * 2 + 1];
        if (row0 == 0) {
            rows[3][0] -= rows[0][0] * u;
            rows[0][0] = 0;
        }
#pragma unroll
        for (int j = 0; j < M; ++j)
            if (row0 + j == len - 1) {
                rows[3][j] -= rows[2][j] * v;
                rows[2][j] = 0;
            }
    }
    T al_r[Sh::NI];
    auto al_store = [&](int j, T v) { al_r[j] = v; };
    SweepOut<T, Sh::NI> o;
    RegLoader<T, Sh::CH, M> ld{rows};
    // row numbers shifted so that the sweep masks nothing: the couplings at the
    // chunk ends were set above
    lane_sweep<T, Sh>(o, ld, row0 + 1, INT64_MAX, al_store);
    const T DLr = o.bL - o.cL * o.SA, ELr = o.cL * o.PI, FLr = o.dL - o.cL * o.SD;
    const T ARr = -(o.aR * o.alp), DRr = o.bR - o.aR * o.g, FRr = o.dR - o.aR * o.del;

    if constexpr (PASS == kChunkReduce) {
        Outer<T> oL, oR;
        segment_outer(o.aL, DLr, ELr, FLr, ARr, DRr, o.cR, FRr, oL, oR);
        wm_wave_merges<T, 64, 1>(oL, oR, lane, 0, lane, 64, false, nullptr);
        // wave blocks: first row from lane 0, last row from lane 63
        __shared__ T sh_w[NW][6];
        if (lane == 0) { sh_w[wave][0] = oL.r; sh_w[wave][1] = oL.p; sh_w[wave][2] = oL.q; }
        if (lane == 63) { sh_w[wave][3] = oR.r; sh_w[wave][4] = oR.p; sh_w[wave][5] = oR.q; }
        __syncthreads();
        if (tid == 0) {
            // fold the wave blocks left to right: (F, E) of the blocks merged so far
            Outer<T> F{sh_w[0][0], sh_w[0][1], sh_w[0][2]}, E{sh_w[0][3], sh_w[0][4], sh_w[0][5]};
            for (int w = 1; w < NW; ++w) {
                const Outer<T> BF{sh_w[w][0], sh_w[w][1], sh_w[w][2]};
                Outer<T> BE{sh_w[w][3], sh_w[w][4], sh_w[w][5]};
                T RB, PB, QB, RA, PA, QA;
                wm_pair_coeffs(E.r, E.p, E.q, BF.r, BF.p, BF.q, RB, PB, QB, RA, PA, QA);
                wm_merge_row(F, true, RB, PB, QB, RA, PA, QA);
                wm_merge_row(BE, false, RB, PB, QB, RA, PA, QA);
                E = BE;
            }
            T *out = p.red + (s * K + k) * 6;
            out[0] = F.r; out[1] = F.p; out[2] = F.q;
            out[3] = E.r; out[4] = E.p; out[5] = E.q;
        }
    } else {
        T xL, xR;
        interface_solve<T, 1, 64, NW, 1, 64 * NW, LANES>(o.aL, DLr, ELr, FLr, ARr, DRr, o.cR, FRr,
                                                         tid, lane, 0, lane, wave, p.iface, 0,
                                                         tri_dyn_lds, xL, xR);
        if (wrow >= len) return;                          // wave-uniform
        T xs[M];
        xs[0] = xL;
        xs[M - 1] = xR;
        T xn = xR;
#pragma unroll
        for (int j = M - 3; j >= 0; --j) {
            xn = (o.de[j] - al_r[j] * xL) - o.ga[j] * xn;
            xs[j + 1] = xn;
        }
        wave_lds_sync();
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            u32x4 t;
#pragma unroll
            for (int v = 0; v < VV; ++v) reinterpret_cast<T *>(&t)[v] = xs[i * VV + v];
            xb[lane * NQ + i] = t;
        }
        wave_lds_sync();
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            tail_chunk_store<T, true>(xb[q * 64 + lane], rr[3], wrow + (int64_t)(q * 64 + lane) * VV,
                                      len);
    }
}

// Pass 2: one workgroup of 64*ceil(K/64) lanes per system (K <= 1024).
template <typename T>
__global__ void __launch_bounds__(1024) chunk_iface_kernel(LargeParams<T> p) {
    __shared__ T sh[1024][6];
    const int t = threadIdx.x, K = p.K;
    const int64_t s = blockIdx.x;
    const bool on = t < K;
    Outer<T> F{0, 0, 0}, E{0, 0, 0};
    if (on) {
        const T *in = p.red + (s * K + t) * 6;
        F = {in[0], in[1], in[2]};
        E = {in[3], in[4], in[5]};
    }
    for (int h = 1; h < K; h <<= 1) {
        if (on) {
            sh[t][0] = F.r; sh[t][1] = F.p; sh[t][2] = F.q;
            sh[t][3] = E.r; sh[t][4] = E.p; sh[t][5] = E.q;
        }
        __syncthreads();
        if (on) {
            const int blk = t / h;
            const int pb = (blk & ~1) * h;                // first chunk of the merged pair
            if (pb + h < K) {                             // the pair has a right block
                const int ia = pb + h - 1, ib = pb + h;   // A's last chunk, B's first chunk
                T RB, PB, QB, RA, PA, QA;
                wm_pair_coeffs(sh[ia][3], sh[ia][4], sh[ia][5], sh[ib][0], sh[ib][1], sh[ib][2],
                               RB, PB, QB, RA, PA, QA);
                const bool inA = (blk & 1) == 0;
                wm_merge_row(F, inA, RB, PB, QB, RA, PA, QA);
                wm_merge_row(E, inA, RB, PB, QB, RA, PA, QA);
            }
        }
        __syncthreads();
    }
    // one block spans the system: x_F = F.r, x_E = E.r
    if (on) { sh[t][0] = F.r; sh[t][1] = E.r; }
    __syncthreads();
    if (on) {
        T *out = p.bnd + (s * K + t) * 2;
        out[0] = (t > 0) ? sh[t - 1][1] : T(0);
        out[1] = (t + 1 < K) ? sh[t + 1][0] : T(0);
    }
}

template <typename T> struct LargeEntry {
    int M, NW, C;
    hipError_t (*launch)(const LargeParams<T> &, hipStream_t);
};

hipError_t launch_large_f32(const LargeParams<float> &p, int shape, hipStream_t st);
hipError_t launch_large_f64(const LargeParams<double> &p, int shape, hipStream_t st);
int large_chunk_rows(int precision, int shape);
int large_shape_count(int precision);

}  // namespace tri
