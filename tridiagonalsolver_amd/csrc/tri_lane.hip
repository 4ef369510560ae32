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
// Thomas over systems of any length in the contiguous (and packed) layout, one
// system per lane, with the reference's memory semantics: c' is written to c and
// d' to d on the way down, x to d on the way up (9 words per unknown).  Each lane
// walks its own rows in chunks of R rows read and written as 16-byte vectors, and
// loads the next chunk while it eliminates the current one, so a system costs
// about 2n/R dependent memory round trips instead of 2n.

template <typename T, int R>
__global__ void __launch_bounds__(64) thomas_rows_kernel(const T *a, const T *b, T *c, T *d,
                                                         int64_t n, int64_t batch, int64_t sstride,
                                                         int32_t bb) {
    constexpr int VV = 16 / (int)sizeof(T);
    constexpr int NV = R / VV;                 // vectors per array and chunk
    static_assert(R % VV == 0, "chunks are whole 16-byte vectors");
    const int64_t s = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (s >= batch) return;                    // no LDS or barriers below
    const __amdgpu_buffer_rsrc_t ra = make_rsrc(a, bb), rb = make_rsrc(b, bb), rc = make_rsrc(c, bb),
                                 rd = make_rsrc(d, bb);
    const int64_t base = s * sstride;
    u32x4 cur[4][NV], nxt[4][NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        cur[0][v] = sys_chunk_load<T>(ra, base, v * VV, n);
        cur[1][v] = sys_chunk_load<T>(rb, base, v * VV, n);
        cur[2][v] = sys_chunk_load<T>(rc, base, v * VV, n);
        cur[3][v] = sys_chunk_load<T>(rd, base, v * VV, n);
    }
    T cp = 0, dp = 0;
    for (int64_t r0 = 0; r0 < n; r0 += R) {
        const bool more = r0 + R < n;
        if (more) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                nxt[0][v] = sys_chunk_load<T>(ra, base, r0 + R + v * VV, n);
                nxt[1][v] = sys_chunk_load<T>(rb, base, r0 + R + v * VV, n);
                nxt[2][v] = sys_chunk_load<T>(rc, base, r0 + R + v * VV, n);
                nxt[3][v] = sys_chunk_load<T>(rd, base, r0 + R + v * VV, n);
            }
        }
        T *A = reinterpret_cast<T *>(cur[0]), *B = reinterpret_cast<T *>(cur[1]);
        T *C = reinterpret_cast<T *>(cur[2]), *D = reinterpret_cast<T *>(cur[3]);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int64_t i = r0 + j;
            if (i < n) {
                const T aj = (i == 0) ? T(0) : A[j];          // a_0 is not part of A
                const T cj = (i == n - 1) ? T(0) : C[j];      // nor is c_{n-1}
                const T k = rcp(B[j] - aj * cp);
                cp = cj * k;
                dp = (D[j] - aj * dp) * k;
                C[j] = cp;
                D[j] = dp;
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            sys_chunk_store<T>(cur[2][v], rc, base, r0 + v * VV, n);
            sys_chunk_store<T>(cur[3][v], rd, base, r0 + v * VV, n);
        }
        if (more) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int v = 0; v < NV; ++v) cur[k][v] = nxt[k][v];
        }
    }
    // back substitution, last chunk first; the last chunk's c', d' are still in cur
    T xn = 0;
    for (int64_t r0 = ((n - 1) / R) * R; r0 >= 0; r0 -= R) {
        const bool more = r0 >= R;
        if (more) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                nxt[2][v] = sys_chunk_load<T>(rc, base, r0 - R + v * VV, n);
                nxt[3][v] = sys_chunk_load<T>(rd, base, r0 - R + v * VV, n);
            }
        }
        T *C = reinterpret_cast<T *>(cur[2]), *D = reinterpret_cast<T *>(cur[3]);
#pragma unroll
        for (int j = R - 1; j >= 0; --j) {
            const int64_t i = r0 + j;
            if (i < n) {
                xn = (i == n - 1) ? D[j] : D[j] - C[j] * xn;
                D[j] = xn;
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) sys_chunk_store<T>(cur[3][v], rd, base, r0 + v * VV, n);
        if (more) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                cur[2][v] = nxt[2][v];
                cur[3][v] = nxt[3][v];
            }
        }
    }
}

// Rows per chunk: 64 bytes of each array.
template <typename T> constexpr int thomas_rows_chunk() { return 64 / (int)sizeof(T); }

template <typename T>
static hipError_t launch_thomas_rows(const T *a, const T *b, T *c, T *d, int64_t n, int64_t batch,
                                     int64_t sstride, int32_t bb, hipStream_t st) {
    constexpr int VV = 16 / (int)sizeof(T);
    auto al = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    if (n < 1 || bb == 0 || sstride % VV != 0 || !al(a) || !al(b) || !al(c) || !al(d))
        return hipErrorInvalidValue;
    if (batch <= 0) return hipSuccess;
    const int64_t blocks = (batch + 63) / 64;
    if (blocks > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((thomas_rows_kernel<T, thomas_rows_chunk<T>()>), dim3((unsigned)blocks),
                       dim3(64), 0, st, a, b, c, d, n, batch, sstride, bb);
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
