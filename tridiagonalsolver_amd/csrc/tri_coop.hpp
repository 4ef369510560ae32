// Cooperative row-block solve: a group of systems whose rows are too many for
// one workgroup is cut into K row blocks, and K workgroups solve the group
// together, each keeping its own block on chip from load to store (5 words of
// HBM traffic per unknown).
//
//   task (g, j) = rows [j*R, (j+1)*R) of the S systems of group g
//     1. the block's rows (interleaved: whole row pieces of S*sizeof(T) bytes;
//        contiguous: whole 16-byte chunks) go to the lanes through LDS; lane
//        sweeps, Wang-Mou merges inside each wave and a fold over the waves give
//        the block's first row F and last row E in outer form relative to
//        U = x of the row before the block and V = x of the row after it
//     2. those 6 values per system are published; the loads of the workgroup's
//        next task are issued; then it waits until the other K-1 blocks of the
//        group have published theirs
//     3. a partial-elimination chain over the K block forms of each system gives
//        U and V of the block, a second chain over the block's waves gives x at
//        every wave boundary, and the lanes back-substitute from registers and
//        store x
//
// Schedule: the grid is P workgroups, all resident at once (P is at most what the
// device holds, and a multiple of K), and workgroup w takes tasks w, w + P,
// w + 2P, ...  The K tasks of a slot are consecutive, so they fall in the same
// round and every member publishes before it waits: with all P workgroups
// resident, every wait ends.  Slot s solves group s of a contiguous batch; in an
// interleaved batch the slots of a round take groups spread over the whole batch
// (coop_group), so that a round reads whole interleaved rows rather than a strip.  Each workgroup keeps the next round's block in
// flight (in registers) while it waits and finishes the current task, so the
// memory system stays busy while groups wait for their slowest member.  Every
// wait is still bounded (spin_limit polls): a workgroup whose partners do not
// publish in time (which needs another kernel to hold CUs that this grid
// counted on) gives up and writes NaN for x of its block, so the kernel always
// ends and a lost wait shows in the result rather than as a wrong finite value.
//
// Publication without L2 writeback or invalidate: the block forms are written
// with agent-scope atomic stores (which write through to the level where all
// XCDs meet), the writing lanes wait until those stores have completed
// (s_waitcnt) before the workgroup barrier, and one lane then stores the solve's
// tag into the block's flag word with an agent-scope atomic store.  The waiting
// side polls the group's flag words with agent-scope atomic loads and reads the
// forms with agent-scope atomic loads too, so no stale copy in an L1 or L2 can
// serve them.  Tagged flags instead of an arrival counter: nothing has to be
// cleared before a solve (a counter needed a reset kernel per solve, 2-6 us).  An
// acquire/release pair at agent scope would also order them, but it compiles to
// an L2 writeback (release) and invalidate (acquire) per workgroup and task,
// which the XCDs' L2s serve one at a time: measured on MI355X, FP32 interleaved
// 4096 x 4096 took 242 us with them and 175 us without (both before the polls
// were spread over separate lines and the schedule was made static).
//
// The reference has no counterpart: its solvers stream one system through one
// warp (Tridiagonal_CUDA/tridiagonal_wm_pcr.cu:17-205) or reduce in global
// memory level by level (tridiagonal_basic.cu:32-164).
#pragma once
#include "tri_coop_sched.hpp"
#include "tri_tile.hpp"

namespace tri {

template <typename T> struct CoopParams {
    const T *a, *b, *c;
    T *d;
    int64_t n, batch;
    int64_t sstride;     // elements between consecutive systems (interleaved: 1)
    int64_t rstride;     // elements between consecutive rows (contiguous: 1)
    int32_t buf_bytes;   // byte extent of each array (below 2 GiB)
    int32_t K;           // row blocks per system
    int32_t ngroups;     // groups of S systems
    int32_t spin_limit;  // polls of a wait before it gives up
    int32_t pipelined;   // 1: coop_pipe_kernel (wait one task later), 0: coop_kernel
    int32_t spread;      // groups per round whose slots are spread over the batch (<= 1: none)
    T *forms;            // [ngroups][K][S][6] block outer forms
    uint64_t *flags;     // [kCoopFlagStride * g + j]: the tag once block j of group g is published
    uint64_t tag;        // this solve's tag: different from every earlier solve's
};

// Publication flags: one 8-byte word per block of a group (up to 64 blocks), the
// group's words in 512 bytes of their own (so that the polls of different groups
// go to different memory channels).  A block is published when its word holds the
// solve's tag.  Tags never repeat (a 64-bit sequence from a random start), so
// words left by earlier solves, or never written, do not read as published and
// the flags need no clearing between solves.
constexpr int kCoopFlagStride = 64;
// s_sleep argument between two polls of a wait (64 clocks per unit).
#ifndef TRI_COOP_POLL_SLEEP
#define TRI_COOP_POLL_SLEEP 8
#endif

// Group solved by slot s: interleaved groups are spread over the batch within a
// round (coop_group); contiguous groups hold one system each and keep slot order.
template <bool IL, typename T>
__device__ __forceinline__ int coop_gi(const CoopParams<T> &p, int s) {
    if constexpr (IL) return coop_group(s, p.ngroups, p.spread);
    else return s;
}

template <typename T, int M, int NW, int S, bool IL> struct CoopShape {
    static constexpr int SW = IL ? 64 / S : 64;      // segments of a system per wave
    static constexpr int R = SW * NW * M;            // rows per block
};

// Wave 0 of a workgroup waits until all K blocks of group g carry the solve's tag:
// lane k polls block k's word, and the wave goes on once every lane has seen it
// (one vote per poll).  Returns 0 if spin_limit polls pass first.
template <typename T>
__device__ __forceinline__ int coop_wait(const CoopParams<T> &p, int g, int lane) {
    const uint64_t *f = p.flags + kCoopFlagStride * g;
    for (int polls = 0;; ++polls) {
        const bool seen =
            lane >= p.K || __hip_atomic_load(f + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == p.tag;
        if (__all(seen)) return 1;
        if (polls >= p.spin_limit) return 0;
        __builtin_amdgcn_s_sleep(TRI_COOP_POLL_SLEEP);
    }
}

// Stages the K block forms of a group (K*S*6 values at fg, block-major) into
// sh_k[system][block][6] with the whole workgroup.  In FP32 each thread issues 8
// loads before it stores the first: the loads are agent-scope (they go past the
// XCD's L2 to where all XCDs meet), and one after another their latencies add up.
// The FP64 kernels have no registers to spare (measured: FP64 65,536 x 256 0.173
// -> 0.182 ms with the batched form) and load one value at a time.  ok == false
// (a wait that gave up): zeros, and no loads.
template <typename T, int S, int NW, int KMAX>
__device__ __forceinline__ void coop_stage_forms(const T *fg, int K, bool ok, int tid,
                                                 T (*sh_k)[KMAX][6]) {
    constexpr int NT = 64 * NW, B = 8;
    const int ne = K * S * 6;
    if constexpr (sizeof(T) != 4) {
        for (int e = tid; e < ne; e += NT) {
            const int k = e / (S * 6), r = e - k * (S * 6);
            sh_k[r / 6][k][r % 6] =
                ok ? __hip_atomic_load(fg + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : T(0);
        }
        return;
    }
    for (int e0 = 0; e0 < ne; e0 += NT * B) {
        T v[B];
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int e = e0 + i * NT + tid;
            v[i] = (ok && e < ne) ? __hip_atomic_load(fg + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : T(0);
        }
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int e = e0 + i * NT + tid;
            if (e < ne) {
                const int k = e / (S * 6), r = e - k * (S * 6);
                sh_k[r / 6][k][r % 6] = v[i];
            }
        }
    }
}

// Folds wave-block outer forms w = 0..nw-1 (first row at [w][0..2], last row at
// [w][3..5]) left to right into the outer form of the whole block.
template <typename T>
__device__ __forceinline__ void fold_blocks(const T (*blk)[6], int nw, Outer<T> &F, Outer<T> &E) {
    F = {blk[0][0], blk[0][1], blk[0][2]};
    E = {blk[0][3], blk[0][4], blk[0][5]};
    for (int w = 1; w < nw; ++w) {
        const Outer<T> BF{blk[w][0], blk[w][1], blk[w][2]};
        Outer<T> BE{blk[w][3], blk[w][4], blk[w][5]};
        T RB, PB, QB, RA, PA, QA;
        wm_pair_coeffs(E.r, E.p, E.q, BF.r, BF.p, BF.q, RB, PB, QB, RA, PA, QA);
        wm_merge_row(F, true, RB, PB, QB, RA, PA, QA);
        wm_merge_row(BE, false, RB, PB, QB, RA, PA, QA);
        E = BE;
    }
}

// Partial elimination over wave-blocks w = 0..nw-1 of one block whose outer
// neighbours are known (x before = U, x after = V): writes u_w = x of the row
// before wave-block w and v_w = x of the row after it.  blk is overwritten.
template <typename T>
__device__ __forceinline__ void chain_with_ends(T (*blk)[6], T (*uv)[2], int nw, T U, T V) {
    T rho = U, sig = 0;                   // x_{E_{w-1}} = rho - sig x_{F_w}
    for (int w = 0; w < nw; ++w) {
        const T rF = blk[w][0], pF = blk[w][1], qF = blk[w][2];
        const T rE = blk[w][3], pE = blk[w][4], qE = blk[w][5];
        const T inv = rcp(T(1) - pF * sig);
        const T rhoF = (rF - pF * rho) * inv;
        const T sigF = qF * inv;
        const T rhoE = rE - pE * (rho - sig * rhoF);
        const T sigE = qE + pE * sig * sigF;
        blk[w][0] = rhoF; blk[w][1] = sigF; blk[w][2] = rhoE; blk[w][3] = sigE;
        rho = rhoE;
        sig = sigE;
    }
    T xFn = V;
    uv[0][0] = U;
    for (int w = nw - 1; w >= 0; --w) {
        const T xE = blk[w][2] - blk[w][3] * xFn;
        const T xF = blk[w][0] - blk[w][1] * xFn;
        uv[w][1] = xFn;
        if (w + 1 < nw) uv[w + 1][0] = xE;
        xFn = xF;
    }
}

// Contiguous-layout stores of x: the wave's 64*M rows as whole 16-byte chunks
// through LDS, rows past the system not written.
template <typename T, int M, bool NT>
__device__ __forceinline__ void contig_block_store(__amdgpu_buffer_rsrc_t rd, int64_t wrow,
                                                   int64_t len, int lane, u32x4 *xb,
                                                   const T (&xs)[M]) {
    constexpr int VV = 16 / (int)sizeof(T);
    constexpr int NQ = M / VV;
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
        tail_chunk_store<T, NT>(xb[q * 64 + lane], rd, wrow + (int64_t)(q * 64 + lane) * VV, len);
}

// block_chain (below) one block at a time: the form the FP64 kernels take.
template <typename T>
__device__ __forceinline__ void block_chain_serial(T (*blk)[6], int K, int j, T &U, T &V) {
    T rho = 0, sig = 0;                   // x_{E_{k-1}} = rho - sig x_{F_k}
    for (int k = 0; k < K; ++k) {
        const T rF = blk[k][0], pF = blk[k][1], qF = blk[k][2];
        const T rE = blk[k][3], pE = blk[k][4], qE = blk[k][5];
        const T inv = rcp(T(1) - pF * sig);
        const T rhoF = (rF - pF * rho) * inv;
        const T sigF = qF * inv;
        blk[k][0] = rhoF; blk[k][1] = sigF;
        blk[k][2] = rho; blk[k][3] = sig;              // relation of x_{E_{k-1}} to x_{F_k}
        rho = rE - pE * (rho - sig * rhoF);
        sig = qE + pE * sig * sigF;
    }
    T xFn = 0;                            // x_{F_{k+1}}
    U = 0;
    V = 0;
    for (int k = K - 1; k >= j; --k) {
        if (k == j) V = xFn;
        const T xF = blk[k][0] - blk[k][1] * xFn;
        if (k == j) U = blk[k][2] - blk[k][3] * xF;
        xFn = xF;
    }
}

// Partial elimination over the K block forms of one system (blk[k][0..5]: first
// row r, p, q then last row r, p, q; the system's ends are zero): returns
// U = x before block j and V = x after it.  blk is overwritten.
template <typename T>
__device__ __forceinline__ void block_chain(T (*blk)[6], int K, int j, T &U, T &V) {
    if constexpr (sizeof(T) != 4) {
        block_chain_serial<T>(blk, K, j, U, V);
        return;
    }
    // FP32: both sweeps take the forms of four blocks from LDS before using the
    // first, so that the LDS latency is paid once per four steps of the dependent
    // chain (FP32 131,072 x 128 contiguous 0.088 -> 0.083 ms).  The FP64 kernels
    // have no registers to spare and keep the step-by-step form.
    constexpr int U4 = 4;
    T rho = 0, sig = 0;                   // x_{E_{k-1}} = rho - sig x_{F_k}
    auto fwd = [&](T (&f)[6], T *out) {
        const T inv = rcp(T(1) - f[1] * sig);
        const T rhoF = (f[0] - f[1] * rho) * inv;
        const T sigF = f[2] * inv;
        out[0] = rhoF; out[1] = sigF;
        out[2] = rho; out[3] = sig;                    // relation of x_{E_{k-1}} to x_{F_k}
        rho = f[3] - f[4] * (rho - sig * rhoF);
        sig = f[5] + f[4] * sig * sigF;
    };
    int k = 0;
    for (; k + U4 <= K; k += U4) {
        T f[U4][6];
#pragma unroll
        for (int i = 0; i < U4; ++i)
#pragma unroll
            for (int q = 0; q < 6; ++q) f[i][q] = blk[k + i][q];
#pragma unroll
        for (int i = 0; i < U4; ++i) fwd(f[i], blk[k + i]);
    }
    for (; k < K; ++k) {
        T f[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) f[q] = blk[k][q];
        fwd(f, blk[k]);
    }
    T xFn = 0;                            // x_{F_{k+1}}
    U = 0;
    V = 0;
    auto bwd = [&](int kk, T rhoF, T sigF, T rin, T sin) {
        if (kk == j) V = xFn;
        const T xF = rhoF - sigF * xFn;
        if (kk == j) U = rin - sin * xF;
        xFn = xF;
    };
    k = K - 1;
    for (; k - (U4 - 1) >= j; k -= U4) {
        T b[U4][4];
#pragma unroll
        for (int i = 0; i < U4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) b[i][q] = blk[k - i][q];
#pragma unroll
        for (int i = 0; i < U4; ++i) bwd(k - i, b[i][0], b[i][1], b[i][2], b[i][3]);
    }
    for (; k >= j; --k) bwd(k, blk[k][0], blk[k][1], blk[k][2], blk[k][3]);
}

// Contiguous blocks: whole 16-byte chunks of the wave's 64*M rows, issued
// unconditionally (the buffer ends at the system's last row, so chunks past it
// read as zero).
template <typename T, int M, bool NT>
__device__ __forceinline__ void contig_block_issue(const __amdgpu_buffer_rsrc_t (&rr)[4], int64_t wrow,
                                                   int lane, u32x4 (&g)[4][M * (int)sizeof(T) / 16]) {
    constexpr int VV = 16 / (int)sizeof(T);
    constexpr int NQ = M / VV;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            g[k][q] = __builtin_amdgcn_raw_buffer_load_b128(
                rr[k], (int)((wrow + (int64_t)(q * 64 + lane) * VV) * (int64_t)sizeof(T)), 0,
                cpol<NT>());
}

template <typename T, int M>
__device__ __forceinline__ void contig_block_transpose(const u32x4 (&g)[4][M * (int)sizeof(T) / 16],
                                                       int lane, u32x4 *xb, T (&rows)[4][M]) {
    constexpr int VV = 16 / (int)sizeof(T);
    constexpr int NQ = M / VV;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k > 0) wave_lds_sync();
#pragma unroll
        for (int q = 0; q < NQ; ++q) xb[q * 64 + lane] = g[k][q];
        wave_lds_sync();
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const u32x4 t = xb[lane * NQ + i];
#pragma unroll
            for (int v = 0; v < VV; ++v) rows[k][i * VV + v] = reinterpret_cast<const T *>(&t)[v];
        }
    }
}

// Buffers of one task: interleaved, the whole arrays; contiguous, from the
// block's first row to the system's last row.
template <typename T, bool IL>
__device__ __forceinline__ void coop_rsrc(const CoopParams<T> &p, int g, int64_t brow,
                                          __amdgpu_buffer_rsrc_t (&rr)[4]) {
    if constexpr (IL) {
        rr[0] = make_rsrc(p.a, p.buf_bytes); rr[1] = make_rsrc(p.b, p.buf_bytes);
        rr[2] = make_rsrc(p.c, p.buf_bytes); rr[3] = make_rsrc(p.d, p.buf_bytes);
    } else {
        const int64_t base = (int64_t)g * p.sstride + brow;
        const int32_t bytes = (int32_t)((p.n - brow) * (int64_t)sizeof(T));
        rr[0] = make_rsrc(p.a + base, bytes); rr[1] = make_rsrc(p.b + base, bytes);
        rr[2] = make_rsrc(p.c + base, bytes); rr[3] = make_rsrc(p.d + base, bytes);
    }
}

// PF: the next task's block is loaded into registers while this one finishes;
// OCC: waves per SIMD the kernel is built for (its register budget).
template <typename T, int M, int NW, int S, bool IL, int KMAX, bool PF, int OCC>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(OCC)))
    coop_kernel(CoopParams<T> p) {
    using Sh = TileShape<T, M, NW, S, IL, !IL, false, false, !IL, true, true>;
    constexpr int SW = IL ? 64 / S : 64;     // segments of a system per wave
    constexpr int R = SW * NW * M;           // rows per block
    constexpr int NQ = M * (int)sizeof(T) / 16;
    static_assert(IL || S == 1, "contiguous groups hold one system");
    static_assert(IL ? (S * (int)sizeof(T)) % 16 == 0 : M % (16 / (int)sizeof(T)) == 0,
                  "blocks move whole 16-byte chunks");
    static_assert(KMAX <= kCoopFlagStride, "a group's flags fit its flag words");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sl = IL ? lane % S : 0;        // system of the lane inside the group
    const int kw = IL ? lane / S : lane;     // segment inside the wave-block
    const int ws = wave;                     // wave-block inside the block
    const int64_t wrow = (int64_t)ws * SW * M;   // first row of the wave-block, in the block
    __shared__ u32x4 xbuf[NW][64 * M * (int)sizeof(T) / 16];
    __shared__ T sh_blk[S][NW][6];
    __shared__ T sh_uv[S][NW][2];
    __shared__ T sh_k[S][KMAX][6];
    __shared__ int s_ok;

    const int K = p.K;
    const int P = (int)gridDim.x;
    const int ntask = p.ngroups * K;
    const int64_t n = p.n;
    int task = (int)blockIdx.x;
    u32x4 g[4][NQ];
    auto issue = [&](int t) {
        const int s0 = t / K, j0 = t - s0 * K;
        const int g0 = coop_gi<IL>(p, s0);
        __amdgpu_buffer_rsrc_t rr[4];
        coop_rsrc<T, IL>(p, g0, (int64_t)j0 * R, rr);
        if constexpr (IL) il_block_issue<T, M, SW, S, true>(rr, (int64_t)j0 * R + wrow, p.rstride,
                                                            (int64_t)g0 * S, lane, g);
        else contig_block_issue<T, M, true>(rr, wrow, lane, g);
    };
    if (PF && task < ntask) issue(task);
    for (; task < ntask; task += P) {
        if constexpr (!PF) issue(task);
        const int si = task / K, j = task - si * K;
        const int gi = coop_gi<IL>(p, si);
        const int64_t brow = (int64_t)j * R;               // first row of the block
        const int64_t row0 = brow + wrow + (int64_t)kw * M;   // first row of the lane
        const int64_t len = IL ? R : ((n - brow < R) ? n - brow : R);
        T rows[4][M];
        if constexpr (IL) {
            il_block_transpose<T, M, S>(g, lane, sl, kw, xbuf[wave], rows);
        } else {
            contig_block_transpose<T, M>(g, lane, xbuf[wave], rows);
#pragma unroll
            for (int q = 0; q < M; ++q)
                if (wrow + kw * M + q >= len) {   // identity rows past the system
                    rows[0][q] = 0; rows[1][q] = 1; rows[2][q] = 0; rows[3][q] = 0;
                }
        }
        // lane sweep; global row numbers drop a_0 and c_{n-1}
        T al_r[Sh::NI];
        SweepOut<T, Sh::NI> o;
        RegLoader<T, Sh::CH, M> ld{rows};
        lane_sweep<T, Sh>(o, ld, row0, n, [&](int q, T x) { al_r[q] = x; });
        const T DLr = o.bL - o.cL * o.SA, ELr = o.cL * o.PI, FLr = o.dL - o.cL * o.SD;
        const T ARr = -(o.aR * o.alp), DRr = o.bR - o.aR * o.g, FRr = o.dR - o.aR * o.del;
        Outer<T> oL, oR;
        segment_outer(o.aL, DLr, ELr, FLr, ARr, DRr, o.cR, FRr, oL, oR);
        wm_wave_merges<T, SW, (IL ? S : 1)>(oL, oR, lane, sl, kw, SW, false, nullptr);
        if (kw == 0) { sh_blk[sl][ws][0] = oL.r; sh_blk[sl][ws][1] = oL.p; sh_blk[sl][ws][2] = oL.q; }
        if (kw == SW - 1) { sh_blk[sl][ws][3] = oR.r; sh_blk[sl][ws][4] = oR.p; sh_blk[sl][ws][5] = oR.q; }
        __syncthreads();

        // publish the block's outer form for every system of the group
        T *fg = p.forms + (int64_t)gi * K * S * 6;
        if (tid < S) {
            Outer<T> F, E;
            fold_blocks<T>(sh_blk[tid], NW, F, E);
            T *f = fg + ((int64_t)j * S + tid) * 6;
            const T vals[6] = {F.r, F.p, F.q, E.r, E.p, E.q};
#pragma unroll
            for (int q = 0; q < 6; ++q)
                __hip_atomic_store(f + q, vals[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // the forms are at the agent-coherent level before the barrier
            __builtin_amdgcn_s_waitcnt(0);
        }
        __syncthreads();
        if (tid == 0)
            __hip_atomic_store(p.flags + kCoopFlagStride * gi + j, p.tag, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        // the next round's block streams in while this group completes
        if (PF && task + P < ntask) issue(task + P);
        if (tid < 64) {
            const int okw = coop_wait(p, gi, tid);
            if (tid == 0) s_ok = okw;
        }
        __syncthreads();
        const bool ok = s_ok != 0;
        coop_stage_forms<T, S, NW, KMAX>(fg, K, ok, tid, sh_k);
        __syncthreads();
        if (tid < S) {
            T U, V;
            block_chain<T>(sh_k[tid], K, j, U, V);
            chain_with_ends<T>(sh_blk[tid], sh_uv[tid], NW, U, V);
        }
        __syncthreads();
        const T u = sh_uv[sl][ws][0], v = sh_uv[sl][ws][1];
        const T xL = oL.r - oL.p * u - oL.q * v;
        const T xR = oR.r - oR.p * u - oR.q * v;
        T xs[M];
        xs[0] = xL;
        xs[M - 1] = xR;
        T xn = xR;
#pragma unroll
        for (int q = M - 3; q >= 0; --q) {
            xn = (o.de[q] - al_r[q] * xL) - o.ga[q] * xn;
            xs[q + 1] = xn;
        }
        if (!ok) {
#pragma unroll
            for (int q = 0; q < M; ++q) xs[q] = __builtin_nan("");
        }
        __amdgpu_buffer_rsrc_t rw[4];
        coop_rsrc<T, IL>(p, gi, brow, rw);
        if constexpr (IL)
            il_block_store<T, M, SW, S, true>(rw[3], brow + wrow, p.rstride, (int64_t)gi * S,
                                              p.batch, lane, sl, kw, xbuf[wave], xs);
        else
            contig_block_store<T, M, true>(rw[3], wrow, len, lane, xbuf[wave], xs);
        __syncthreads();                           // LDS of this task free before the next
    }
}

// Back-substitution of one task once its group has published: waits for the K
// block forms of group gi, chains them to U, V of block j, chains the wave-blocks
// (blk: this task's wave-block forms, overwritten) and stores x of the block.
template <typename T, int M, int NW, int S, bool IL, int KMAX, int NI>
__device__ __forceinline__ void coop_finish(const CoopParams<T> &p, int gi, int j, T (*blk)[NW][6],
                                            T (*sh_uv)[NW][2], T (*sh_k)[KMAX][6], int *s_ok,
                                            u32x4 *xb, int tid, int lane, int sl, int kw, int ws,
                                            int64_t wrow, const Outer<T> &oL, const Outer<T> &oR,
                                            const T (&de)[NI], const T (&ga)[NI],
                                            const T (&al)[NI]) {
    constexpr int SW = IL ? 64 / S : 64;
    constexpr int R = SW * NW * M;
    const int K = p.K;
    const int64_t n = p.n;
    const int64_t brow = (int64_t)j * R;
    const int64_t len = IL ? R : ((n - brow < R) ? n - brow : R);
    const T *fg = p.forms + (int64_t)gi * K * S * 6;
    if (tid < 64) {
        const int ok = coop_wait(p, gi, tid);
        if (tid == 0) *s_ok = ok;
    }
    __syncthreads();
    const bool ok = *s_ok != 0;
    coop_stage_forms<T, S, NW, KMAX>(fg, K, ok, tid, sh_k);
    __syncthreads();
    if (tid < S) {
        T U, V;
        block_chain<T>(sh_k[tid], K, j, U, V);
        chain_with_ends<T>(blk[tid], sh_uv[tid], NW, U, V);
    }
    __syncthreads();
    const T u = sh_uv[sl][ws][0], v = sh_uv[sl][ws][1];
    const T xL = oL.r - oL.p * u - oL.q * v;
    const T xR = oR.r - oR.p * u - oR.q * v;
    T xs[M];
    xs[0] = xL;
    xs[M - 1] = xR;
    T xn = xR;
#pragma unroll
    for (int q = M - 3; q >= 0; --q) {
        xn = (de[q] - al[q] * xL) - ga[q] * xn;
        xs[q + 1] = xn;
    }
    if (!ok) {
#pragma unroll
        for (int q = 0; q < M; ++q) xs[q] = __builtin_nan("");
    }
    __amdgpu_buffer_rsrc_t rw[4];
    coop_rsrc<T, IL>(p, gi, brow, rw);
    if constexpr (IL)
        il_block_store<T, M, SW, S, true>(rw[3], brow + wrow, p.rstride, (int64_t)gi * S, p.batch,
                                          lane, sl, kw, xb, xs);
    else
        contig_block_store<T, M, true>(rw[3], wrow, len, lane, xb, xs);
    __syncthreads();                           // LDS of this task free
}

// The same solve with the wait taken one task later: a workgroup reduces and
// publishes task t + P before it waits for the group of task t, so the time the
// group's slowest member needs is covered by a whole task's loads and arithmetic
// instead of idling.  The lane state of the pending task (its outer forms and the
// sweep coefficients of its rows, 3(M-2) + 6 values per lane) stays in registers.
// Every wait for task t still comes after the workgroup has published task t + P,
// and every member of a group publishes it before waiting on anything later, so
// with all P workgroups resident every wait ends.
template <typename T, int M, int NW, int S, bool IL, int KMAX>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(2)))
    coop_pipe_kernel(CoopParams<T> p) {
    using Sh = TileShape<T, M, NW, S, IL, !IL, false, false, !IL, true, true>;
    constexpr int SW = IL ? 64 / S : 64;
    constexpr int R = SW * NW * M;
    constexpr int NQ = M * (int)sizeof(T) / 16;
    constexpr int NI = Sh::NI;
    static_assert(IL || S == 1, "contiguous groups hold one system");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sl = IL ? lane % S : 0;
    const int kw = IL ? lane / S : lane;
    const int ws = wave;
    const int64_t wrow = (int64_t)ws * SW * M;
    __shared__ u32x4 xbuf[NW][64 * M * (int)sizeof(T) / 16];
    __shared__ T sh_blk[2][S][NW][6];
    __shared__ T sh_uv[S][NW][2];
    __shared__ T sh_k[S][KMAX][6];
    __shared__ int s_ok;

    const int K = p.K;
    const int P = (int)gridDim.x;
    const int ntask = p.ngroups * K;
    const int64_t n = p.n;
    int task = (int)blockIdx.x;
    u32x4 g[4][NQ];
    if (task < ntask) {
        const int s0 = task / K, j0 = task - s0 * K;
        const int g0 = coop_gi<IL>(p, s0);
        __amdgpu_buffer_rsrc_t rr[4];
        coop_rsrc<T, IL>(p, g0, (int64_t)j0 * R, rr);
        if constexpr (IL) il_block_issue<T, M, SW, S, true>(rr, (int64_t)j0 * R + wrow, p.rstride,
                                                            (int64_t)g0 * S, lane, g);
        else contig_block_issue<T, M, true>(rr, wrow, lane, g);
    }
    // the pending task: published, not yet finished
    int pend = -1, par = 0;
    Outer<T> poL{}, poR{};
    T pde[NI], pga[NI], pal[NI];
    for (; task < ntask; task += P) {
        const int si = task / K, j = task - si * K;
        const int gi = coop_gi<IL>(p, si);
        const int64_t brow = (int64_t)j * R;
        const int64_t row0 = brow + wrow + (int64_t)kw * M;
        const int64_t len = IL ? R : ((n - brow < R) ? n - brow : R);
        T rows[4][M];
        if constexpr (IL) {
            il_block_transpose<T, M, S>(g, lane, sl, kw, xbuf[wave], rows);
        } else {
            contig_block_transpose<T, M>(g, lane, xbuf[wave], rows);
#pragma unroll
            for (int q = 0; q < M; ++q)
                if (wrow + kw * M + q >= len) {
                    rows[0][q] = 0; rows[1][q] = 1; rows[2][q] = 0; rows[3][q] = 0;
                }
        }
        T al_r[NI];
        SweepOut<T, NI> o;
        RegLoader<T, Sh::CH, M> ld{rows};
        lane_sweep<T, Sh>(o, ld, row0, n, [&](int q, T x) { al_r[q] = x; });
        const T DLr = o.bL - o.cL * o.SA, ELr = o.cL * o.PI, FLr = o.dL - o.cL * o.SD;
        const T ARr = -(o.aR * o.alp), DRr = o.bR - o.aR * o.g, FRr = o.dR - o.aR * o.del;
        Outer<T> oL, oR;
        segment_outer(o.aL, DLr, ELr, FLr, ARr, DRr, o.cR, FRr, oL, oR);
        wm_wave_merges<T, SW, (IL ? S : 1)>(oL, oR, lane, sl, kw, SW, false, nullptr);
        if (kw == 0) {
            sh_blk[par][sl][ws][0] = oL.r; sh_blk[par][sl][ws][1] = oL.p; sh_blk[par][sl][ws][2] = oL.q;
        }
        if (kw == SW - 1) {
            sh_blk[par][sl][ws][3] = oR.r; sh_blk[par][sl][ws][4] = oR.p; sh_blk[par][sl][ws][5] = oR.q;
        }
        __syncthreads();
        T *fg = p.forms + (int64_t)gi * K * S * 6;
        if (tid < S) {
            Outer<T> F, E;
            fold_blocks<T>(sh_blk[par][tid], NW, F, E);
            T *f = fg + ((int64_t)j * S + tid) * 6;
            const T vals[6] = {F.r, F.p, F.q, E.r, E.p, E.q};
#pragma unroll
            for (int q = 0; q < 6; ++q)
                __hip_atomic_store(f + q, vals[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_s_waitcnt(0);
        }
        __syncthreads();
        if (tid == 0)
            __hip_atomic_store(p.flags + kCoopFlagStride * gi + j, p.tag, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        const int next = task + P;
        if (next < ntask) {
            const int s1 = next / K, j1 = next - s1 * K;
            const int g1 = coop_gi<IL>(p, s1);
            __amdgpu_buffer_rsrc_t rr[4];
            coop_rsrc<T, IL>(p, g1, (int64_t)j1 * R, rr);
            if constexpr (IL) il_block_issue<T, M, SW, S, true>(rr, (int64_t)j1 * R + wrow,
                                                                p.rstride, (int64_t)g1 * S, lane, g);
            else contig_block_issue<T, M, true>(rr, wrow, lane, g);
        }
        if (pend >= 0) {
            const int ps = pend / K;
            coop_finish<T, M, NW, S, IL, KMAX, NI>(p, coop_gi<IL>(p, ps),
                                                   pend - ps * K, sh_blk[par ^ 1], sh_uv,
                                                   sh_k, &s_ok, xbuf[wave], tid, lane, sl, kw, ws,
                                                   wrow, poL, poR, pde, pga, pal);
        }
        pend = task;
        par ^= 1;
        poL = oL;
        poR = oR;
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            pde[q] = o.de[q];
            pga[q] = o.ga[q];
            pal[q] = al_r[q];
        }
    }
    if (pend >= 0) {
        const int ps = pend / K;
        coop_finish<T, M, NW, S, IL, KMAX, NI>(p, coop_gi<IL>(p, ps), pend - ps * K,
                                               sh_blk[par ^ 1], sh_uv, sh_k,
                                               &s_ok, xbuf[wave], tid, lane, sl, kw, ws, wrow, poL,
                                               poR, pde, pga, pal);
    }
}


template <typename T> struct CoopEntry {
    int M, NW, S;
    bool IL;
    int R;       // rows per block
    int kmax;    // most blocks per system (LDS for the block chain)
    bool pipe;   // coop_pipe_kernel by default
    hipError_t (*launch)(const CoopParams<T> &, int grid, hipStream_t);
    int (*max_grid)(bool pipelined);   // co-resident workgroups of the kernel (0: unknown)
};
template <typename T> struct CoopTable {
    const CoopEntry<T> *entries;
    int count;
};
CoopTable<float> coop_table_f32();
CoopTable<double> coop_table_f64();

}  // namespace tri
