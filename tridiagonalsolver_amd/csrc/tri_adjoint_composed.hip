// Adjoint of systems the fused adjoint kernels do not hold (long systems, or
// interleaved systems above their capacity), composed from the plain solver:
//
//   A^T has sub-diagonal c_{i-1}, diagonal b_i and super-diagonal a_{i+1}, so
//   lambda = A^{-T} g is an ordinary tridiagonal solve on shifted coefficients,
//   and the coefficient gradients follow element-wise:
//     ga_i = -lambda_i x_{i-1},  gb_i = -lambda_i x_i,  gc_i = -lambda_i x_{i+1},
//     gd_i = lambda_i            (ga_0 = gc_{n-1} = 0).
//
// The shifted coefficients are written into the ga / gc outputs and g into gd,
// the solve runs in place on gd, and the gradient kernel then overwrites ga, gb,
// gc: no memory beyond the caller's outputs.  Element (s, i) is at s*n + i
// (contiguous) or i*batch + s (interleaved).
#include <algorithm>

#include "tri_host.hpp"

namespace tri {

template <typename T>
__global__ void __launch_bounds__(256) adj_shift_kernel(const T *__restrict__ a,
                                                        const T *__restrict__ c,
                                                        const T *__restrict__ g,
                                                        const T *__restrict__ d, T *__restrict__ ap,
                                                        T *__restrict__ cp, T *__restrict__ lam,
                                                        T *__restrict__ xo, int64_t n,
                                                        int64_t batch, bool il) {
    const int64_t total = n * batch;
    const int64_t step = il ? batch : 1;           // distance between rows i and i+1
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * 256) {
        const int64_t i = il ? e / batch : e % n;
        ap[e] = (i >= 1) ? c[e - step] : T(0);     // sub-diagonal of A^T
        cp[e] = (i <= n - 2) ? a[e + step] : T(0); // super-diagonal of A^T
        lam[e] = g[e];
        if (d) xo[e] = d[e];                       // fused: x is solved in place from d
    }
}

template <typename T>
__global__ void __launch_bounds__(256) adj_grad_kernel(const T *__restrict__ x,
                                                       const T *__restrict__ lam,
                                                       T *__restrict__ ga, T *__restrict__ gb,
                                                       T *__restrict__ gc, int64_t n,
                                                       int64_t batch, bool il) {
    const int64_t total = n * batch;
    const int64_t step = il ? batch : 1;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * 256) {
        const int64_t i = il ? e / batch : e % n;
        const T l = lam[e];
        ga[e] = (i >= 1) ? -l * x[e - step] : T(0);
        gb[e] = -l * x[e];
        gc[e] = (i <= n - 2) ? -l * x[e + step] : T(0);
    }
}

static unsigned elementwise_blocks(int64_t total) {
    return (unsigned)std::min<int64_t>((total + 255) / 256, 256 * 64);
}

template <typename T>
static hipError_t launch_adj_shift(const T *a, const T *c, const T *g, const T *d, T *ap, T *cp,
                                   T *lam, T *xo, int64_t n, int64_t batch, bool il,
                                   hipStream_t st) {
    const int64_t total = n * batch;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL((adj_shift_kernel<T>), dim3(elementwise_blocks(total)), dim3(256), 0, st, a,
                       c, g, d, ap, cp, lam, xo, n, batch, il);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_adj_grad(const T *x, const T *lam, T *ga, T *gb, T *gc, int64_t n,
                                  int64_t batch, bool il, hipStream_t st) {
    const int64_t total = n * batch;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL((adj_grad_kernel<T>), dim3(elementwise_blocks(total)), dim3(256), 0, st, x,
                       lam, ga, gb, gc, n, batch, il);
    return hipGetLastError();
}

hipError_t launch_adj_shift_f32(const float *a, const float *c, const float *g, const float *d,
                                float *ap, float *cp, float *lam, float *xo, int64_t n,
                                int64_t batch, bool il, hipStream_t st) {
    return launch_adj_shift<float>(a, c, g, d, ap, cp, lam, xo, n, batch, il, st);
}
hipError_t launch_adj_shift_f64(const double *a, const double *c, const double *g,
                                const double *d, double *ap, double *cp, double *lam, double *xo,
                                int64_t n, int64_t batch, bool il, hipStream_t st) {
    return launch_adj_shift<double>(a, c, g, d, ap, cp, lam, xo, n, batch, il, st);
}
hipError_t launch_adj_grad_f32(const float *x, const float *lam, float *ga, float *gb, float *gc,
                               int64_t n, int64_t batch, bool il, hipStream_t st) {
    return launch_adj_grad<float>(x, lam, ga, gb, gc, n, batch, il, st);
}
hipError_t launch_adj_grad_f64(const double *x, const double *lam, double *ga, double *gb,
                               double *gc, int64_t n, int64_t batch, bool il, hipStream_t st) {
    return launch_adj_grad<double>(x, lam, ga, gb, gc, n, batch, il, st);
}

}  // namespace tri
