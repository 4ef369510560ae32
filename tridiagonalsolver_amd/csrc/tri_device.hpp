// Device primitives shared by every solver kernel (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tri {

// ---------------------------------------------------------------- reciprocal
// Every division of the solvers is a reciprocal followed by products (one pivot
// reciprocal serves several quotients).  The division policy is fixed when the
// library is built, as the reference fixes its division macros at compile time
// (Tridiagonal_CUDA/intrin.h:23-87):
//   TRIDIAG_DIV_MODE 0 "refined" (default): hardware reciprocal (v_rcp_f32 /
//     v_rcp_f64) refined by Newton-Raphson, y' = y + y(1 - x y): one step in FP32,
//     two in FP64, which brings the estimate to working precision;
//   1 "ieee": the correctly rounded reciprocal, 1/x by IEEE division;
//   2 "fast": the hardware estimate alone in FP32 (1 ulp), one Newton step in FP64.
// Zero pivots give inf/NaN in every mode, which propagate to the solution as they
// would with IEEE division.
#ifndef TRIDIAG_DIV_MODE
#define TRIDIAG_DIV_MODE 0
#endif
constexpr int kDivMode = TRIDIAG_DIV_MODE;
static_assert(kDivMode >= 0 && kDivMode <= 2, "TRIDIAG_DIV_MODE must be 0, 1 or 2");

__device__ __forceinline__ float rcp(float x) {
    if constexpr (kDivMode == 1) return 1.0f / x;
    const float y = __builtin_amdgcn_rcpf(x);
    if constexpr (kDivMode == 2) return y;
    return __builtin_fmaf(y, __builtin_fmaf(-x, y, 1.0f), y);
}
__device__ __forceinline__ double rcp(double x) {
    if constexpr (kDivMode == 1) return 1.0 / x;
    double y = __builtin_amdgcn_rcp(x);
    y = __builtin_fma(y, __builtin_fma(-x, y, 1.0), y);
    if constexpr (kDivMode == 2) return y;
    return __builtin_fma(y, __builtin_fma(-x, y, 1.0), y);
}

// ---------------------------------------------------------------- cross-lane
// Read `v` from lane `src` of the same wave (ds_bpermute: byte address = 4*lane).
__device__ __forceinline__ float lane_read(float v, int src) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ double lane_read(double v, int src) {
    int2 p = *reinterpret_cast<int2 *>(&v);
    p.x = __builtin_amdgcn_ds_bpermute(src << 2, p.x);
    p.y = __builtin_amdgcn_ds_bpermute(src << 2, p.y);
    return *reinterpret_cast<double *>(&p);
}

// Orders LDS traffic between the lanes of one wave: every LDS write issued
// before it is visible to every lane's LDS reads after it, and the compiler may
// not move LDS accesses across it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; static constexpr int N = 4; };
template <> struct Vec16<double> { using type = double2; static constexpr int N = 2; };

// ---------------------------------------------------------------- outer form
// A block of consecutive rows is in "outer form" when every row i of it reads
//     x_i = r_i - p_i * u - q_i * v,
// with u = x of the row just before the block and v = x of the row just after
// it.  A lane-segment, a wave-block and a WG-block all carry their first and
// last rows in this form.
template <typename T> struct Outer {
    T r, p, q;
};

// Merge step of Wang & Mou (1991): blocks A and B are adjacent (A left of B).
// Given A's last row (rA,pA,qA) and B's first row (rB,pB,qB) in outer form, a
// row of A (inA) or of B is rewritten in outer form of the merged block.
template <typename T>
__device__ __forceinline__ void wm_merge_row(Outer<T> &x, bool inA, T RB, T PB, T QB, T RA, T PA,
                                             T QA) {
    if (inA) {
        // v_A = x_{B first} = RB - PB u - QB v
        x.r = x.r - x.q * RB;
        x.p = x.p - x.q * PB;
        x.q = -x.q * QB;
    } else {
        // u_B = x_{A last} = RA - PA u - QA v
        x.r = x.r - x.p * RA;
        x.q = x.q - x.p * QA;
        x.p = -x.p * PA;
    }
}

template <typename T>
__device__ __forceinline__ void wm_pair_coeffs(T rA, T pA, T qA, T rB, T pB, T qB, T &RB, T &PB,
                                               T &QB, T &RA, T &PA, T &QA) {
    // 2x2 system for (x_{A last}, x_{B first}):
    //   x_Al + qA x_Bf = rA - pA u ;  pB x_Al + x_Bf = rB - qB v
    T inv = rcp(T(1) - qA * pB);
    RB = (rB - pB * rA) * inv;
    PB = -(pB * pA) * inv;
    QB = qB * inv;
    RA = (rA - qA * rB) * inv;
    PA = pA * inv;
    QA = -(qA * qB) * inv;
}

}  // namespace tri
