// Short systems, one system per lane, held in registers.
//
// The Thomas recurrence of the oracle (Tridiagonal_AVX512/tridiagonal_basic.cpp:6-26):
//   c'_0 = c_0 / b_0,  d'_0 = d_0 / b_0,
//   c'_i = c_i / (b_i - a_i c'_{i-1}),  d'_i = (d_i - a_i d'_{i-1}) / (b_i - a_i c'_{i-1}),
//   x_{n-1} = d'_{n-1},  x_i = d'_i - c'_i x_{i+1},
// with c' and d' kept in VGPRs (the plain Thomas kernel writes them back to
// memory), so a solve moves 5 words per unknown.  For systems of up to N rows:
// the contiguous layout reads each lane's rows as 16-byte pieces, the interleaved
// layout reads one row of 64 consecutive systems per load instruction.
#include <stdlib.h>

#include "tri_host.hpp"

namespace tri {

// Forward and backward Thomas sweeps over rows held in registers (rows past n
// are identity rows).  x replaces D.
template <typename T, int N>
__device__ __forceinline__ void lane_thomas_sweeps(T (&A)[N], T (&B)[N], T (&C)[N], T (&D)[N],
                                                   int64_t n) {
    T cp = 0, dp = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const T aj = (j == 0) ? T(0) : A[j];              // a_0 is not part of A
        const T cj = (j == n - 1) ? T(0) : C[j];          // nor is c_{n-1}
        const T k = rcp(B[j] - aj * cp);
        cp = cj * k;
        dp = (D[j] - aj * dp) * k;
        C[j] = cp;
        D[j] = dp;
    }
    T xn = 0;
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        xn = D[j] - C[j] * xn;
        D[j] = xn;
    }
}

// Interleaved layout: lane = system, one row of 64 consecutive systems per load.
template <typename T, int N>
__global__ void __launch_bounds__(256) lane_thomas_il_kernel(const T *a, const T *b, const T *c,
                                                             T *d, int64_t n, int64_t batch,
                                                             int64_t rstride, int32_t buf_bytes) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= batch) return;                       // no workgroup barriers below
    const __amdgpu_buffer_rsrc_t ra = make_rsrc(a, buf_bytes), rb = make_rsrc(b, buf_bytes),
                                 rc = make_rsrc(c, buf_bytes), rd = make_rsrc(d, buf_bytes);
    T A[N], B[N], C[N], D[N];
    const int vo = (int)(s * (int64_t)sizeof(T));
    const int rowb = (int)(rstride * (int64_t)sizeof(T));
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const bool in = j < n;
        A[j] = in ? bload<T, true>(ra, vo, j * rowb) : T(0);
        B[j] = in ? bload<T, true>(rb, vo, j * rowb) : T(1);
        C[j] = in ? bload<T, true>(rc, vo, j * rowb) : T(0);
        D[j] = in ? bload<T, true>(rd, vo, j * rowb) : T(0);
    }
    lane_thomas_sweeps<T, N>(A, B, C, D, n);
#pragma unroll
    for (int j = 0; j < N; ++j)
        if (j < n) bstore<T, true>(D[j], rd, vo, j * rowb);
}

// Contiguous layout (system stride n): the 64 systems of a wave are 64*n consecutive
// elements, read as whole 16-byte chunks by consecutive lanes and handed to the lanes
// through LDS, where system s occupies an odd stride of P >= n elements so that the
// lanes' row reads hit distinct banks.
template <typename T, int N>
__global__ void __launch_bounds__(256) lane_thomas_c_kernel(const T *a, const T *b, const T *c,
                                                            T *d, int64_t n, int64_t batch) {
    constexpr int VV = 16 / (int)sizeof(T);
    constexpr int NQ = (N + VV - 1) / VV;          // chunks per lane at n = N
    constexpr int PMAX = N | 1;
    __shared__ T lds[4][64 * PMAX];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t sys0 = ((int64_t)blockIdx.x * 4 + wave) * 64;   // first system of the wave
    if (sys0 >= batch) return;                                    // wave-uniform, no barriers
    const int64_t nsw = (batch - sys0 < 64) ? batch - sys0 : 64;   // systems in this wave
    const int64_t lim = nsw * n;                                    // elements of the wave block
    const int P = (int)(n | 1);
    const float inv_n = 1.0f / (float)n;
    const int32_t nb = (int32_t)(lim * (int64_t)sizeof(T));
    const int64_t base = sys0 * n;
    const __amdgpu_buffer_rsrc_t rr[4] = {make_rsrc(a + base, nb), make_rsrc(b + base, nb),
                                          make_rsrc(c + base, nb), make_rsrc(d + base, nb)};
    T *L = lds[wave];
    // element e of the block -> LDS slot (e / n) * P + e % n
    auto slot = [&](int e) {
        const int q = (int)(((float)e + 0.5f) * inv_n);
        return q * P + (e - q * (int)n);
    };
    T R[4][N];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        u32x4 g[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = (q * 64 + lane) * VV;
            g[q] = (q * 64 * VV < 64 * n) ? tail_chunk_load<T, true>(rr[k], e, lim)
                                          : u32x4{0u, 0u, 0u, 0u};
        }
        if (k > 0) wave_lds_sync();
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = (q * 64 + lane) * VV;
#pragma unroll
            for (int v = 0; v < VV; ++v)
                if (e + v < lim) L[slot(e + v)] = reinterpret_cast<const T *>(&g[q])[v];
        }
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < N; ++j) R[k][j] = (j < n) ? L[lane * P + j] : (k == 1 ? T(1) : T(0));
    }
    lane_thomas_sweeps<T, N>(R[0], R[1], R[2], R[3], n);
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < N; ++j)
        if (j < n) L[lane * P + j] = R[3][j];
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = (q * 64 + lane) * VV;
        if (e >= lim) continue;
        u32x4 t = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int v = 0; v < VV; ++v)
            if (e + v < lim) reinterpret_cast<T *>(&t)[v] = L[slot(e + v)];
        tail_chunk_store<T, true>(t, rr[3], e, lim);
    }
}

template <typename T, int N>
static hipError_t launch_lane(const T *a, const T *b, const T *c, T *d, int64_t n, int64_t batch,
                              int64_t rstride, bool il, int32_t bb, hipStream_t st) {
    const int64_t blocks = (batch + 255) / 256;
    if (blocks > 0x7fffffff) return hipErrorInvalidValue;
    if (il)
        hipLaunchKernelGGL((lane_thomas_il_kernel<T, N>), dim3((unsigned)blocks), dim3(256), 0, st,
                           a, b, c, d, n, batch, rstride, bb);
    else
        hipLaunchKernelGGL((lane_thomas_c_kernel<T, N>), dim3((unsigned)blocks), dim3(256), 0, st,
                           a, b, c, d, n, batch);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_lane_thomas(const T *a, const T *b, const T *c, T *d, int64_t n,
                                     int64_t batch, int64_t sstride, int64_t rstride, bool il,
                                     int32_t bb, hipStream_t st) {
    if (n < 1 || n > lane_thomas_max_rows(sizeof(T) == 4 ? 0 : 1)) return hipErrorInvalidValue;
    if (!il && sstride != n) return hipErrorInvalidValue;   // contiguous systems only
    if (batch <= 0) return hipSuccess;
#define TRI_LANE(NN) \
    if (n <= NN) return launch_lane<T, NN>(a, b, c, d, n, batch, rstride, il, bb, st);
    TRI_LANE(8)
    TRI_LANE(16)
    if constexpr (sizeof(T) == 4) { TRI_LANE(32) }
#undef TRI_LANE
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------- longer systems
// Thomas over systems of any length in the contiguous (and packed) layout, with
// the reference's memory semantics: c' is written to c and d' to d on the way
// down, x to d on the way up (9 words per unknown).  The recurrence of a system
// is serial, so its speed is set by how many of its rows are in flight: a lane
// that walked its system alone would wait one memory round trip (~1 us) per few
// rows.  Here L lanes share a system.  A step covers SR = L * 16/w rows: lane t of
// the system's group loads the t-th 16-byte vector of each array, the group's
// vectors pass through LDS to all L lanes, and every lane of the group runs the
// same recurrence over the SR rows, keeping and storing only its own vector of
// c', d' (and later x).  Loads run DF steps ahead on the way down and DB steps
// ahead on the way up.  Each row is stored and read back by the same lane.

template <typename T, int L, int DF, int DB>
__global__ void __launch_bounds__(64) thomas_team_kernel(const T *a, const T *b, T *c, T *d,
                                                         int64_t n, int64_t batch, int64_t sstride,
                                                         int32_t bb) {
    constexpr int VV = 16 / (int)sizeof(T);    // rows per 16-byte vector
    constexpr int G = 64 / L;                  // systems per wave
    constexpr int SR = L * VV;                 // rows per step
    static_assert(64 % L == 0, "a wave holds whole groups");
    __shared__ u32x4 xs[G][4][L];
    const int lane = threadIdx.x, g = lane / L, t = lane % L;
    const int64_t s = (int64_t)blockIdx.x * G + g;
    const bool valid = s < batch;
    const int64_t nl = valid ? n : 0;          // rows this lane may touch (none when idle)
    const int64_t base = valid ? s * sstride : 0;
    const __amdgpu_buffer_rsrc_t rr[4] = {make_rsrc(a, bb), make_rsrc(b, bb), make_rsrc(c, bb),
                                          make_rsrc(d, bb)};
    const int64_t steps = (n + SR - 1) / SR;
    const int64_t own = (int64_t)t * VV;       // first row of the lane's vector in a step
    // Vector loads take no bounds branch: a branch ends in a copy that waits for the
    // newest load and so drains the prefetch ring.  sstride is a multiple of VV and
    // at least n, so the vector holding row n-1 ends inside the system's stride;
    // vectors outside [0, n), and idle lanes, read the array's first vector instead,
    // which is never used.
    auto vload = [&](int r, int64_t e) -> u32x4 {
        const int off = ((uint64_t)e < (uint64_t)nl) ? (int)((base + e) * (int64_t)sizeof(T)) : 0;
        return __builtin_amdgcn_raw_buffer_load_b128(rr[r], off, 0, cpol<false>());
    };

    // ---- forward elimination
    // Slots are issued in the order they are consumed (the scheduler would otherwise
    // reorder the independent loads, and waiting for the first slot would then wait
    // for all of them).
    u32x4 ring[DF][4];
#pragma unroll
    for (int k = 0; k < DF; ++k) {
#pragma unroll
        for (int r = 0; r < 4; ++r) ring[k][r] = vload(r, k * SR + own);
        __builtin_amdgcn_sched_barrier(0);
    }
    T cp = 0, dp = 0;
    for (int64_t q0 = 0; q0 < steps; q0 += DF) {
#pragma unroll
        for (int k = 0; k < DF; ++k) {
            // Every step of the group hands over its slot and refills it, also steps
            // past the end: a refill skipped on some path would leave the compiler
            // unable to count the loads in flight, and it would then wait for all of
            // them.  Only the elimination and the stores are conditional.
            const int64_t q = q0 + k;
            const int64_t r0 = q * SR;
            wave_lds_sync();                   // the group has read the previous step
#pragma unroll
            for (int r = 0; r < 4; ++r) xs[g][r][t] = ring[k][r];
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < 4; ++r) ring[k][r] = vload(r, (q + DF) * SR + own);
            if (q >= steps) continue;
            u32x4 oc = {0u, 0u, 0u, 0u}, od = oc;
            // GUARD: the step holds row 0, row n-1 or rows past n.  Interior steps run
            // without branches, so the LDS reads of the next vector and the selects
            // fill the latency of the reciprocal chain.
            auto eliminate = [&](auto guard) {
                constexpr bool GUARD = decltype(guard)::value;
#pragma unroll
                for (int u = 0; u < L; ++u) {
                    const u32x4 va = xs[g][0][u], vb = xs[g][1][u];
                    u32x4 vc = xs[g][2][u], vd = xs[g][3][u];
                    const T *A = reinterpret_cast<const T *>(&va);
                    const T *B = reinterpret_cast<const T *>(&vb);
                    T *C = reinterpret_cast<T *>(&vc), *D = reinterpret_cast<T *>(&vd);
#pragma unroll
                    for (int v = 0; v < VV; ++v) {
                        const int64_t i = r0 + u * VV + v;
                        if (!GUARD || i < n) {
                            const T aj = (GUARD && i == 0) ? T(0) : A[v];      // a_0 is not part of A
                            const T cj = (GUARD && i == n - 1) ? T(0) : C[v];  // nor is c_{n-1}
                            const T k2 = rcp(B[v] - aj * cp);
                            cp = cj * k2;
                            dp = (D[v] - aj * dp) * k2;
                            C[v] = cp;
                            D[v] = dp;
                        }
                    }
                    if (u == t) { oc = vc; od = vd; }
                }
            };
            if (r0 > 0 && r0 + SR < n)
                eliminate(std::integral_constant<bool, false>{});
            else
                eliminate(std::integral_constant<bool, true>{});
            sys_chunk_store<T>(oc, rr[2], base, r0 + own, nl);
            sys_chunk_store<T>(od, rr[3], base, r0 + own, nl);
        }
    }

    // ---- back substitution, last step first (c', d' read back by the lanes that
    // stored them)
    u32x4 back[DB][2];
#pragma unroll
    for (int k = 0; k < DB; ++k) {
#pragma unroll
        for (int r = 0; r < 2; ++r) back[k][r] = vload(2 + r, (steps - 1 - k) * SR + own);
        __builtin_amdgcn_sched_barrier(0);
    }
    T xn = 0;
    for (int64_t q0 = steps - 1; q0 >= 0; q0 -= DB) {
#pragma unroll
        for (int k = 0; k < DB; ++k) {
            const int64_t q = q0 - k;          // unconditional hand-over and refill, as above
            const int64_t r0 = q * SR;
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < 2; ++r) xs[g][2 + r][t] = back[k][r];
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < 2; ++r) back[k][r] = vload(2 + r, (q - DB) * SR + own);
            if (q < 0) continue;
            u32x4 ox = {0u, 0u, 0u, 0u};
            // GUARD: the step holds row n-1 or rows past n
            auto substitute = [&](auto guard) {
                constexpr bool GUARD = decltype(guard)::value;
#pragma unroll
                for (int u = L - 1; u >= 0; --u) {
                    const u32x4 vc = xs[g][2][u];
                    u32x4 vd = xs[g][3][u];
                    const T *C = reinterpret_cast<const T *>(&vc);
                    T *D = reinterpret_cast<T *>(&vd);
#pragma unroll
                    for (int v = VV - 1; v >= 0; --v) {
                        const int64_t i = r0 + u * VV + v;
                        if (!GUARD || i < n) {
                            xn = (GUARD && i == n - 1) ? D[v] : D[v] - C[v] * xn;
                            D[v] = xn;
                        }
                    }
                    if (u == t) ox = vd;
                }
            };
            if (r0 + SR < n)
                substitute(std::integral_constant<bool, false>{});
            else
                substitute(std::integral_constant<bool, true>{});
            sys_chunk_store<T>(ox, rr[3], base, r0 + own, nl);
        }
    }
}

// ---------------------------------------------------------------- Thomas by wave scans
// The same recurrence and memory semantics (c' into c and d' into d on the way
// down, x into d on the way up: 9 words per unknown), one system per wave, with
// the serial chain evaluated by parallel prefix over the wave's 64 lanes.  A step
// covers SR = 64 * 16/w rows; lane t holds the t-th 16-byte vector (VV rows) of
// each array.
//   c'_j = c_j / (b_j - a_j c'_{j-1}) is the linear-fractional map with matrix
//     [[0, c_j], [-a_j, b_j]] applied to c'_{j-1}.  Each lane composes the maps of
//     its VV rows; an exclusive scan of the composed 2x2 matrices over the lanes
//     gives every lane c' of the row before its vector from the step's incoming
//     c' with one division.  The lane then runs the recurrence over its own VV
//     rows, which gives each row's pivot t_j = b_j - a_j c'_{j-1} and 1/t_j.
//     The matrices are scaled by powers of two (exact) after every product:
//     the map does not change, and no entry overflows.
//   d'_j = (d_j - a_j d'_{j-1}) / t_j is affine in d'_{j-1}; x_j = d'_j - c'_j x_{j+1}
//     is affine in x_{j+1}.  Both take the same exclusive scan of composed maps
//     (forward, then backward over the lanes).
// The dependence carried from one step to the next is the step's last c' and d'
// (and, going back, its first x): a few operations per SR rows instead of a
// division chain per row.  Rounding differs from the row-by-row recurrence; the
// results agree with it to the solver's tolerance.
template <typename T> struct Mob {
    T m00, m01, m10, m11;   // c_out = (m00 c_in + m01) / (m10 c_in + m11)
};

__device__ __forceinline__ int frexp_exp(float x) { return __builtin_amdgcn_frexp_expf(x); }
__device__ __forceinline__ int frexp_exp(double x) { return __builtin_amdgcn_frexp_exp(x); }
__device__ __forceinline__ float ldexp_t(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
__device__ __forceinline__ double ldexp_t(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }

// B applied after A, scaled so that the largest entry lies in [0.5, 1).
template <typename T>
__device__ __forceinline__ Mob<T> mob_after(const Mob<T> &B, const Mob<T> &A) {
    Mob<T> r{B.m00 * A.m00 + B.m01 * A.m10, B.m00 * A.m01 + B.m01 * A.m11,
             B.m10 * A.m00 + B.m11 * A.m10, B.m10 * A.m01 + B.m11 * A.m11};
    const T mx = fmax(fmax(fabs(r.m00), fabs(r.m01)), fmax(fabs(r.m10), fabs(r.m11)));
    const int e = __builtin_isfinite(mx) ? frexp_exp(mx) : 0;   // 0 for mx == 0
    r.m00 = ldexp_t(r.m00, -e);
    r.m01 = ldexp_t(r.m01, -e);
    r.m10 = ldexp_t(r.m10, -e);
    r.m11 = ldexp_t(r.m11, -e);
    return r;
}

// Affine map y = p + q x.  B applied after A.
template <typename T> struct Aff {
    T p, q;
};
template <typename T> __device__ __forceinline__ Aff<T> aff_after(const Aff<T> &B, const Aff<T> &A) {
    return {B.p + B.q * A.p, B.q * A.q};
}

// Moves between lanes.  DPP (data-parallel primitives: the source lane is picked by
// a modifier on the VALU move, a few cycles) or ds_bpermute (LDS crossbar, tens of
// cycles).  dpp_mov<CTRL, ROWS>(old, x): x of the lane that the DPP control CTRL
// names, or old where that lane lies outside the row or the wave, and in rows not
// in the mask ROWS.
template <int CTRL, int ROWS> __device__ __forceinline__ float dpp_mov(float old, float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, x), CTRL,
                                                                 ROWS, 0xf, false));
}
template <int CTRL, int ROWS> __device__ __forceinline__ double dpp_mov(double old, double x) {
    const uint64_t o = __builtin_bit_cast(uint64_t, old), v = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)o, (int)(uint32_t)v,
                                                              CTRL, ROWS, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(
        (int)(uint32_t)(o >> 32), (int)(uint32_t)(v >> 32), CTRL, ROWS, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float read_lane(float x, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}
__device__ __forceinline__ double read_lane(double x, int l) {
    const uint64_t v = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// DPP controls of GFX9: row_shr:d 0x110+d, row_shl:d 0x100+d (inside rows of 16
// lanes), wave_shr:1 0x138, wave_shl:1 0x130, row_bcast:15 0x142 (lane 15 of each
// row into the next row), row_bcast:31 0x143 (lane 31 into rows 2 and 3).
template <int CTRL, int ROWS, typename T> __device__ __forceinline__ Mob<T> mob_dpp(const Mob<T> &x) {
    return {dpp_mov<CTRL, ROWS>(T(1), x.m00), dpp_mov<CTRL, ROWS>(T(0), x.m01),
            dpp_mov<CTRL, ROWS>(T(0), x.m10), dpp_mov<CTRL, ROWS>(T(1), x.m11)};
}
template <int CTRL, int ROWS, typename T> __device__ __forceinline__ Aff<T> aff_dpp(const Aff<T> &x) {
    return {dpp_mov<CTRL, ROWS>(T(0), x.p), dpp_mov<CTRL, ROWS>(T(1), x.q)};
}

// Scans over the wave.  Inclusive up: lane t ends with X_t o ... o X_0; exclusive
// up: X_(t-1) o ... o X_0 (identity at lane 0).  Inclusive down: X_t o ... o X_63
// (maps applied from the last lane to the first); exclusive down: X_(t+1) o ... o
// X_63 (identity at lane 63).  Lanes without a partner compose with the identity,
// so the scans have no lane-dependent branches.
template <bool DPP> struct WaveScan;
template <> struct WaveScan<true> {
    template <typename T> __device__ static Mob<T> up(Mob<T> x, int) {
        x = mob_after(x, mob_dpp<0x111, 0xf>(x));
        x = mob_after(x, mob_dpp<0x112, 0xf>(x));
        x = mob_after(x, mob_dpp<0x114, 0xf>(x));
        x = mob_after(x, mob_dpp<0x118, 0xf>(x));
        x = mob_after(x, mob_dpp<0x142, 0xa>(x));
        return mob_after(x, mob_dpp<0x143, 0xc>(x));
    }
    template <typename T> __device__ static Mob<T> excl_up(const Mob<T> &x, int) {
        return mob_dpp<0x138, 0xf>(x);
    }
    template <typename T> __device__ static Aff<T> up(Aff<T> x, int) {
        x = aff_after(x, aff_dpp<0x111, 0xf>(x));
        x = aff_after(x, aff_dpp<0x112, 0xf>(x));
        x = aff_after(x, aff_dpp<0x114, 0xf>(x));
        x = aff_after(x, aff_dpp<0x118, 0xf>(x));
        x = aff_after(x, aff_dpp<0x142, 0xa>(x));
        return aff_after(x, aff_dpp<0x143, 0xc>(x));
    }
    template <typename T> __device__ static Aff<T> excl_up(const Aff<T> &x, int) {
        return aff_dpp<0x138, 0xf>(x);
    }
    template <typename T> __device__ static Aff<T> down(Aff<T> x, int lane) {
        x = aff_after(x, aff_dpp<0x101, 0xf>(x));
        x = aff_after(x, aff_dpp<0x102, 0xf>(x));
        x = aff_after(x, aff_dpp<0x104, 0xf>(x));
        x = aff_after(x, aff_dpp<0x108, 0xf>(x));
        // lane 16r now holds row r's composite; rows below take those of the rows above
        const Aff<T> t3{read_lane(x.p, 48), read_lane(x.q, 48)};
        const Aff<T> t2 = aff_after(Aff<T>{read_lane(x.p, 32), read_lane(x.q, 32)}, t3);
        const Aff<T> t1 = aff_after(Aff<T>{read_lane(x.p, 16), read_lane(x.q, 16)}, t2);
        const int row = lane >> 4;
        const Aff<T> y = row == 0 ? t1 : row == 1 ? t2 : row == 2 ? t3 : Aff<T>{T(0), T(1)};
        return aff_after(x, y);
    }
    template <typename T> __device__ static Aff<T> excl_down(const Aff<T> &x, int) {
        return aff_dpp<0x130, 0xf>(x);
    }
    template <typename T> __device__ static T at(T x, int l) { return read_lane(x, l); }
};
template <> struct WaveScan<false> {
    template <typename T> __device__ static T sh_up(T v, int d) { return __shfl_up(v, (unsigned)d, 64); }
    template <typename T> __device__ static T sh_dn(T v, int d) { return __shfl_down(v, (unsigned)d, 64); }
    template <typename T> __device__ static Mob<T> up(Mob<T> x, int lane) {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const Mob<T> y{sh_up(x.m00, d), sh_up(x.m01, d), sh_up(x.m10, d), sh_up(x.m11, d)};
            if (lane >= d) x = mob_after(x, y);
        }
        return x;
    }
    template <typename T> __device__ static Mob<T> excl_up(const Mob<T> &x, int lane) {
        const Mob<T> y{sh_up(x.m00, 1), sh_up(x.m01, 1), sh_up(x.m10, 1), sh_up(x.m11, 1)};
        return lane == 0 ? Mob<T>{T(1), T(0), T(0), T(1)} : y;
    }
    template <typename T> __device__ static Aff<T> up(Aff<T> x, int lane) {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const Aff<T> y{sh_up(x.p, d), sh_up(x.q, d)};
            if (lane >= d) x = aff_after(x, y);
        }
        return x;
    }
    template <typename T> __device__ static Aff<T> excl_up(const Aff<T> &x, int lane) {
        const Aff<T> y{sh_up(x.p, 1), sh_up(x.q, 1)};
        return lane == 0 ? Aff<T>{T(0), T(1)} : y;
    }
    template <typename T> __device__ static Aff<T> down(Aff<T> x, int lane) {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const Aff<T> y{sh_dn(x.p, d), sh_dn(x.q, d)};
            if (lane + d < 64) x = aff_after(x, y);
        }
        return x;
    }
    template <typename T> __device__ static Aff<T> excl_down(const Aff<T> &x, int lane) {
        const Aff<T> y{sh_dn(x.p, 1), sh_dn(x.q, 1)};
        return lane == 63 ? Aff<T>{T(0), T(1)} : y;
    }
    template <typename T> __device__ static T at(T x, int l) { return __shfl(x, l, 64); }
};

template <typename T, int NV, int DF, int DB, bool DPP>
__global__ void __launch_bounds__(256) thomas_scan_kernel(const T *a, const T *b, T *c, T *d,
                                                          int64_t n, int64_t batch, int64_t sstride,
                                                          int32_t bb) {
    constexpr int VV = 16 / (int)sizeof(T);   // rows per 16-byte vector
    constexpr int R = NV * VV;                 // rows per lane and step
    constexpr int SR = 64 * R;                 // rows per step
    using W = WaveScan<DPP>;
    const int lane = threadIdx.x & 63;
    const int64_t s = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (s >= batch) return;                    // whole waves; no workgroup barriers below
    const int64_t base = s * sstride;
    const __amdgpu_buffer_rsrc_t rr[4] = {make_rsrc(a, bb), make_rsrc(b, bb), make_rsrc(c, bb),
                                          make_rsrc(d, bb)};
    const int64_t steps = (n + SR - 1) / SR;
    const int64_t own = (int64_t)lane * R;
    // unguarded vector loads, as in thomas_team_kernel: sstride is a multiple of VV
    // and at least n; vectors outside the system read the array's first vector
    auto vload = [&](int r, int64_t e) -> u32x4 {
        const int off = ((uint64_t)e < (uint64_t)n) ? (int)((base + e) * (int64_t)sizeof(T)) : 0;
        return __builtin_amdgcn_raw_buffer_load_b128(rr[r], off, 0, cpol<false>());
    };

    // ---- forward: c' and d'
    u32x4 ring[DF][4][NV];
#pragma unroll
    for (int k = 0; k < DF; ++k) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int v = 0; v < NV; ++v) ring[k][r][v] = vload(r, k * SR + own + v * VV);
        __builtin_amdgcn_sched_barrier(0);
    }
    T cp_in = 0, dp_in = 0;                    // c', d' of the row before the step
    for (int64_t q0 = 0; q0 < steps; q0 += DF) {
#pragma unroll
        for (int k = 0; k < DF; ++k) {
            const int64_t q = q0 + k;
            const int64_t r0 = q * SR + own;   // first row of the lane's rows
            u32x4 cur[4][NV];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    cur[r][v] = ring[k][r][v];
                    ring[k][r][v] = vload(r, (q + DF) * SR + own + v * VV);
                }
            if (q >= steps) continue;          // wave-uniform
            T A[R], B[R], C[R], D[R];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int64_t i = r0 + j;
                const bool in = i < n;
                const int v = j / VV, e = j % VV;
                A[j] = (in && i > 0) ? reinterpret_cast<const T *>(&cur[0][v])[e] : T(0);
                B[j] = in ? reinterpret_cast<const T *>(&cur[1][v])[e] : T(1);
                C[j] = (in && i < n - 1) ? reinterpret_cast<const T *>(&cur[2][v])[e] : T(0);
                D[j] = in ? reinterpret_cast<const T *>(&cur[3][v])[e] : T(0);
            }
            Mob<T> P{T(0), C[0], -A[0], B[0]};
#pragma unroll
            for (int j = 1; j < R; ++j) P = mob_after(Mob<T>{T(0), C[j], -A[j], B[j]}, P);
            const Mob<T> E = W::excl_up(W::up(P, lane), lane);
            T cprev = (E.m00 * cp_in + E.m01) * rcp(E.m10 * cp_in + E.m11);
            T al[R], ga[R];
            u32x4 oc[NV];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const T it = rcp(B[j] - A[j] * cprev);
                cprev = C[j] * it;
                reinterpret_cast<T *>(&oc[j / VV])[j % VV] = cprev;
                al[j] = D[j] * it;
                ga[j] = -(A[j] * it);
            }
            Aff<T> L{al[0], ga[0]};
#pragma unroll
            for (int j = 1; j < R; ++j) L = aff_after(Aff<T>{al[j], ga[j]}, L);
            const Aff<T> F = W::excl_up(W::up(L, lane), lane);
            T dprev = F.p + F.q * dp_in;
            u32x4 od[NV];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                dprev = al[j] + ga[j] * dprev;
                reinterpret_cast<T *>(&od[j / VV])[j % VV] = dprev;
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                sys_chunk_store<T>(oc[v], rr[2], base, r0 + v * VV, n);
                sys_chunk_store<T>(od[v], rr[3], base, r0 + v * VV, n);
            }
            cp_in = W::at(cprev, 63);
            dp_in = W::at(dprev, 63);
        }
    }

    // ---- backward: x_j = d'_j - c'_j x_{j+1}, last step first (each lane reads back
    // the vectors it stored)
    u32x4 back[DB][2][NV];
#pragma unroll
    for (int k = 0; k < DB; ++k) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int v = 0; v < NV; ++v)
                back[k][r][v] = vload(2 + r, (steps - 1 - k) * SR + own + v * VV);
        __builtin_amdgcn_sched_barrier(0);
    }
    T x_in = 0;                                // x of the row after the step
    for (int64_t q0 = steps - 1; q0 >= 0; q0 -= DB) {
#pragma unroll
        for (int k = 0; k < DB; ++k) {
            const int64_t q = q0 - k;
            const int64_t r0 = q * SR + own;
            u32x4 cur[2][NV];
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    cur[r][v] = back[k][r][v];
                    back[k][r][v] = vload(2 + r, (q - DB) * SR + own + v * VV);
                }
            if (q < 0) continue;
            T Cp[R], Dp[R];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const bool in = r0 + j < n;
                Cp[j] = in ? reinterpret_cast<const T *>(&cur[0][j / VV])[j % VV] : T(0);
                Dp[j] = in ? reinterpret_cast<const T *>(&cur[1][j / VV])[j % VV] : T(0);
            }
            // the lane's map from x after its rows to x of its first row
            Aff<T> L{Dp[R - 1], -Cp[R - 1]};
#pragma unroll
            for (int j = R - 2; j >= 0; --j) L = aff_after(Aff<T>{Dp[j], -Cp[j]}, L);
            const Aff<T> F = W::excl_down(W::down(L, lane), lane);
            T xn = F.p + F.q * x_in;
            u32x4 ox[NV];
#pragma unroll
            for (int j = R - 1; j >= 0; --j) {
                xn = Dp[j] - Cp[j] * xn;
                reinterpret_cast<T *>(&ox[j / VV])[j % VV] = xn;
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) sys_chunk_store<T>(ox[v], rr[3], base, r0 + v * VV, n);
            x_in = W::at(xn, 0);
        }
    }
}

// $TRIDIAG_THOMAS_SCAN: 0 keeps the 16-lane team kernel, 2 moves the scan kernel's
// values between lanes through ds_bpermute instead of DPP; $TRIDIAG_THOMAS_NV fixes
// the 16-byte vectors per lane and step (1 or 2).
static int thomas_scan_mode() {
    static const int v = [] {
        const char *e = getenv("TRIDIAG_THOMAS_SCAN");
        return (e && (e[0] == '0' || e[0] == '2')) ? e[0] - '0' : 1;
    }();
    return v;
}
// Two vectors per lane halve the steps, and with them the scans on the chain: the
// better choice while the batch leaves most SIMDs without a second wave.  Measured
// on MI355X (ms, one / two vectors): FP64 1024 x 1024 0.0179 / 0.0168, FP64 512 x
// 1024 0.0111 / 0.0103, FP64 65536 x 256 0.687 / 0.539; at 4096 x 4096 one vector
// keeps more loads in flight per register: FP64 0.203 / 0.211, FP32 0.103 / 0.111.
static int thomas_scan_nv(int64_t batch) {
    static const int env = [] {
        const char *e = getenv("TRIDIAG_THOMAS_NV");
        return (e && (e[0] == '1' || e[0] == '2')) ? e[0] - '0' : 0;
    }();
    return env ? env : (batch <= 2048 ? 2 : 1);
}

template <typename T>
static hipError_t launch_thomas_rows(const T *a, const T *b, T *c, T *d, int64_t n, int64_t batch,
                                     int64_t sstride, int32_t bb, hipStream_t st) {
    constexpr int VV = 16 / (int)sizeof(T);
    constexpr int L = 16;                      // lanes per system
    auto al = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    // the kernel's unguarded vector loads rely on sstride >= n, a multiple of VV
    if (n < 1 || bb == 0 || sstride < n || sstride % VV != 0 || !al(a) || !al(b) || !al(c) ||
        !al(d))
        return hipErrorInvalidValue;
    if (batch <= 0) return hipSuccess;
    const int64_t blocks = (batch + 64 / L - 1) / (64 / L);
    if (blocks > 0x7fffffff) return hipErrorInvalidValue;
    if (thomas_scan_mode() != 0) {
        const int64_t wg = (batch + 3) / 4;
        if (wg > 0x7fffffff) return hipErrorInvalidValue;
        const dim3 grid((unsigned)wg), block(256);
        if (thomas_scan_mode() == 2)
            hipLaunchKernelGGL((thomas_scan_kernel<T, 1, 4, 8, false>), grid, block, 0, st, a, b,
                               c, d, n, batch, sstride, bb);
        else if (thomas_scan_nv(batch) == 2)
            hipLaunchKernelGGL((thomas_scan_kernel<T, 2, 2, 4, true>), grid, block, 0, st, a, b, c,
                               d, n, batch, sstride, bb);
        else
            hipLaunchKernelGGL((thomas_scan_kernel<T, 1, 4, 8, true>), grid, block, 0, st, a, b, c,
                               d, n, batch, sstride, bb);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((thomas_team_kernel<T, L, 4, 8>), dim3((unsigned)blocks), dim3(64), 0, st,
                       a, b, c, d, n, batch, sstride, bb);
    return hipGetLastError();
}

hipError_t launch_thomas_rows_f32(const float *a, const float *b, float *c, float *d, int64_t n,
                                  int64_t batch, int64_t sstride, int32_t bb, hipStream_t st) {
    return launch_thomas_rows<float>(a, b, c, d, n, batch, sstride, bb, st);
}
hipError_t launch_thomas_rows_f64(const double *a, const double *b, double *c, double *d,
                                  int64_t n, int64_t batch, int64_t sstride, int32_t bb,
                                  hipStream_t st) {
    return launch_thomas_rows<double>(a, b, c, d, n, batch, sstride, bb, st);
}

int lane_thomas_max_rows(int precision) { return precision == 0 ? 32 : 16; }

hipError_t launch_lane_thomas_f32(const float *a, const float *b, const float *c, float *d,
                                  int64_t n, int64_t batch, int64_t sstride, int64_t rstride,
                                  bool il, int32_t bb, hipStream_t st) {
    return launch_lane_thomas<float>(a, b, c, d, n, batch, sstride, rstride, il, bb, st);
}
hipError_t launch_lane_thomas_f64(const double *a, const double *b, const double *c, double *d,
                                  int64_t n, int64_t batch, int64_t sstride, int64_t rstride,
                                  bool il, int32_t bb, hipStream_t st) {
    return launch_lane_thomas<double>(a, b, c, d, n, batch, sstride, rstride, il, bb, st);
}

}  // namespace tri
