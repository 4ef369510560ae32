// Layout conversion between the three coefficient layouts of the solver.
//
// Element (coefficient k, system s, row i) lives at
//   layout 0 (contiguous):   ptr_k[s*n + i]
//   layout 1 (BPLG packed):  buf[s*4n + k*n + i]   (Tridiagonal_CUDA/utils.cu:274-283)
//   layout 2 (interleaved):  ptr_k[i*batch + s]
// so every conversion is, per coefficient, either a row-pitched copy of a
// rows x cols matrix (0 <-> 1, and same-layout copies) or its transpose
// (0/1 <-> 2).  Both kernels are HBM-bound streams: 2 * sizeof(T) bytes per
// element, read once and written once.
#include <algorithm>

#include "tri_host.hpp"

namespace tri {

// 16-byte clang vector of T (the nontemporal builtins take these, not HIP's vector structs).
template <typename T> struct RawVec16 {
    typedef T type __attribute__((ext_vector_type(16 / sizeof(T))));
};

// Transpose of TR x TC tiles through LDS: dst[c * ld_dst + r] = src[r * ld_src + c].
// Global loads and stores are 16-byte vectors along the contiguous direction of
// each side (VEC) or single elements; vectors that cross the matrix edge fall
// back to element accesses.  blockIdx.x is the tile, blockIdx.y the coefficient.
// TR = 64 rows, TC = 512 bytes of T: the loads read 512-byte row pieces.  Measured
// on MI355X (four 4096^2 arrays, contiguous to interleaved, ms): FP32 64 x 64 0.111,
// 64 x 128 0.102, 128 x 64 0.111, 128 x 128 0.125; FP64 64 x 64 0.190, 32 x 64
// 0.232, 64 x 32 0.227, 32 x 32 0.270.
template <typename T> constexpr int transpose_rows() { return 64; }
template <typename T> constexpr int transpose_cols() { return 512 / (int)sizeof(T); }
template <typename T, bool VEC>
__global__ void __launch_bounds__(256) layout_transpose_kernel(LayoutMove<T> p) {
    constexpr int TR = transpose_rows<T>(), TC = transpose_cols<T>();
    constexpr int V = VEC ? 16 / (int)sizeof(T) : 1;
    constexpr int VPR = TC / V;  // vectors per tile row (loads)
    constexpr int VPC = TR / V;  // vectors per tile column (stores)
    using vec_t = typename RawVec16<T>::type;
    __shared__ T tile[TR][TC + 1];
    const T *__restrict__ src = p.src[blockIdx.y];
    T *__restrict__ dst = p.dst[blockIdx.y];
    const int64_t t = blockIdx.x;
    const int64_t r0 = (t / p.col_tiles) * TR, c0 = (t % p.col_tiles) * TC;
#pragma unroll
    for (int k = threadIdx.x; k < TR * VPR; k += 256) {
        const int rr = k / VPR, cc = (k % VPR) * V;
        const int64_t r = r0 + rr, c = c0 + cc;
        if (r >= p.rows) continue;
        const T *s = src + r * p.ld_src + c;
        if (VEC && c + V <= p.cols) {
            const vec_t v = __builtin_nontemporal_load(reinterpret_cast<const vec_t *>(s));
#pragma unroll
            for (int j = 0; j < V; ++j) tile[rr][cc + j] = v[j];
        } else {
#pragma unroll
            for (int j = 0; j < V; ++j)
                if (c + j < p.cols) tile[rr][cc + j] = s[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = threadIdx.x; k < TC * VPC; k += 256) {
        const int cc = k / VPC, rr = (k % VPC) * V;
        const int64_t c = c0 + cc, r = r0 + rr;
        if (c >= p.cols) continue;
        T *o = dst + c * p.ld_dst + r;
        if (VEC && r + V <= p.rows) {
            vec_t v;
#pragma unroll
            for (int j = 0; j < V; ++j) v[j] = tile[rr + j][cc];
            __builtin_nontemporal_store(v, reinterpret_cast<vec_t *>(o));
        } else {
#pragma unroll
            for (int j = 0; j < V; ++j)
                if (r + j < p.rows) o[j] = tile[rr + j][cc];
        }
    }
}

// Row-pitched copy: dst[r * ld_dst + c] = src[r * ld_src + c] for a rows x cols
// matrix, one 16-byte vector (VEC) or element per thread and step of a
// grid-stride loop over all rows.
template <typename T, bool VEC>
__global__ void __launch_bounds__(256) layout_copy_kernel(LayoutMove<T> p, int64_t vpr,
                                                          int64_t total) {
    constexpr int V = VEC ? 16 / (int)sizeof(T) : 1;
    using vec_t = typename RawVec16<T>::type;
    const T *__restrict__ src = p.src[blockIdx.y];
    T *__restrict__ dst = p.dst[blockIdx.y];
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < total; g += stride) {
        const int64_t r = g / vpr, c = (g - r * vpr) * V;
        const T *s = src + r * p.ld_src + c;
        T *o = dst + r * p.ld_dst + c;
        if (VEC && c + V <= p.cols) {
            const vec_t v = __builtin_nontemporal_load(reinterpret_cast<const vec_t *>(s));
            __builtin_nontemporal_store(v, reinterpret_cast<vec_t *>(o));
        } else {
#pragma unroll
            for (int j = 0; j < V; ++j)
                if (c + j < p.cols) o[j] = s[j];
        }
    }
}

template <typename T> static bool aligned_16(const void *q) {
    return (reinterpret_cast<uintptr_t>(q) & 15) == 0;
}

template <typename T>
static hipError_t launch_layout_move(LayoutMove<T> p, int count, bool transpose, hipStream_t st) {
    if (count < 1 || count > 4 || p.rows < 1 || p.cols < 1) return hipErrorInvalidValue;
    constexpr int V = 16 / (int)sizeof(T);
    bool vec = (p.ld_src % V) == 0 && (p.ld_dst % V) == 0;
    for (int k = 0; k < count; ++k) vec = vec && aligned_16<T>(p.src[k]) && aligned_16<T>(p.dst[k]);
    if (transpose) {
        constexpr int TR = transpose_rows<T>(), TC = transpose_cols<T>();
        p.col_tiles = (p.cols + TC - 1) / TC;
        const int64_t tiles = ((p.rows + TR - 1) / TR) * p.col_tiles;
        if (tiles > 0x7fffffff) return hipErrorInvalidValue;
        const dim3 grid((unsigned)tiles, (unsigned)count);
        if (vec) hipLaunchKernelGGL((layout_transpose_kernel<T, true>), grid, dim3(256), 0, st, p);
        else hipLaunchKernelGGL((layout_transpose_kernel<T, false>), grid, dim3(256), 0, st, p);
    } else {
        const int64_t vpr = vec ? (p.cols + V - 1) / V : p.cols;
        const int64_t total = p.rows * vpr;
        // enough workgroups to fill the chip several times over; the loop takes the rest
        const int64_t blocks = std::min<int64_t>((total + 255) / 256, 256 * 64);
        const dim3 grid((unsigned)blocks, (unsigned)count);
        if (vec) hipLaunchKernelGGL((layout_copy_kernel<T, true>), grid, dim3(256), 0, st, p, vpr, total);
        else hipLaunchKernelGGL((layout_copy_kernel<T, false>), grid, dim3(256), 0, st, p, vpr, total);
    }
    return hipGetLastError();
}

hipError_t launch_layout_move_f32(const LayoutMove<float> &p, int count, bool transpose,
                                  hipStream_t st) {
    return launch_layout_move<float>(p, count, transpose, st);
}
hipError_t launch_layout_move_f64(const LayoutMove<double> &p, int count, bool transpose,
                                  hipStream_t st) {
    return launch_layout_move<double>(p, count, transpose, st);
}

}  // namespace tri
