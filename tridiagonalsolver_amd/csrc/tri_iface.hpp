// Interface solvers of the tile kernels: after the lane sweep every lane holds
// the two boundary rows (L, R) of its segment; the rows of all segments of a
// system form a tridiagonal "interface" system of 2P rows, ordered
// L_0, R_0, L_1, R_1, ...  (L_k couples R_{k-1}, L_k, R_k; R_k couples L_k, R_k,
// L_{k+1}).  These routines solve it with parallel cyclic reduction (PCR),
// PCR followed by per-lane Thomas on the remaining strided subsystems, or
// cyclic reduction (CR), all through LDS.  The Wang-Mou merge + pGE chain lives
// in tri_tile.hpp.
#pragma once
#include "tri_device.hpp"

namespace tri {

enum Iface : int {
    kIfaceWM = 0,       // Wang-Mou merges over the whole wave-block, pGE chain over waves
    kIfaceWMHalf = 1,   // merges over half-wave blocks, pGE chain over 2x as many blocks
    kIfacePCR = 2,      // PCR over the whole interface system
    kIfacePCRThomas = 3,// PCR with strides inside a wave-block, then Thomas across waves
    kIfaceCR = 4,       // cyclic reduction
};

// An interface-solver code is the family (Iface) in bits 0-3 and, for
// kIfacePCRThomas, the PCR depth in bits 4-7: the PCR steps taken before Thomas
// runs along the remaining subsystems.  d steps leave 2^d interleaved subsystems
// of the 2P interface rows (the first step pairs each segment's two rows, each
// further one doubles the segment stride); depth 0 is the family's default, one
// step per doubling up to the wave's segment stride.
__host__ __device__ constexpr int iface_family(int code) { return code & 15; }
__host__ __device__ constexpr int iface_pcr_depth(int code) { return (code >> 4) & 15; }

template <typename T> struct Row4 {
    T a, b, c, d;
};

template <typename T> __device__ __forceinline__ Row4<T> ident_row() { return {T(0), T(1), T(0), T(0)}; }

// One PCR elimination: row x with its neighbours lo (above) and hi (below) at the
// current stride; the result couples x to the rows one stride beyond lo and hi.
template <typename T>
__device__ __forceinline__ Row4<T> pcr_combine(const Row4<T> &x, const Row4<T> &lo,
                                               const Row4<T> &hi) {
    const T k1 = x.a * rcp(lo.b);
    const T k2 = x.c * rcp(hi.b);
    Row4<T> y;
    y.a = -(lo.a * k1);
    y.c = -(hi.c * k2);
    y.b = x.b - lo.c * k1 - hi.a * k2;
    y.d = x.d - lo.d * k1 - hi.d * k2;
    return y;
}

// LDS view of the interface systems: S systems x N rows, four coefficient arrays.
template <typename T> struct IfaceLds {
    T *a, *b, *c, *d;
    int N;
    __device__ __forceinline__ Row4<T> get(int sys, int r) const {
        if (r < 0 || r >= N) return ident_row<T>();
        const int i = sys * N + r;
        return {a[i], b[i], c[i], d[i]};
    }
    __device__ __forceinline__ void put(int sys, int r, const Row4<T> &x) const {
        const int i = sys * N + r;
        a[i] = x.a; b[i] = x.b; c[i] = x.c; d[i] = x.d;
    }
};

template <typename T> __device__ __forceinline__ IfaceLds<T> iface_lds(void *base, int S, int N) {
    T *p = reinterpret_cast<T *>(base);
    const int SN = S * N;
    return {p, p + SN, p + 2 * SN, p + 3 * SN, N};
}

// Bytes of dynamic LDS the LDS-based interface solvers need for a workgroup of
// `lanes` lanes (S*P == lanes, 2 rows per lane, 4 coefficients per row).
template <typename T> constexpr int iface_lds_bytes(int lanes) { return 8 * lanes * (int)sizeof(T); }

// PCR on the interface rows.  `seg` = segment index of the lane inside its
// system (0..P-1), `sys` = system inside the workgroup.  steps_stride_limit:
// PCR stops once the segment stride reaches it (P for full PCR); the rows then
// couple only to rows `limit` segments away and the caller finishes them.
template <typename T>
__device__ __forceinline__ void pcr_interface(const IfaceLds<T> &L, int sys, int seg, int P,
                                              int stride_limit, Row4<T> &rl, Row4<T> &rr) {
    // step 0 (row stride 1): L_k with R_{k-1} and R_k; R_k with L_k and L_{k+1}
    L.put(sys, 2 * seg, rl);
    L.put(sys, 2 * seg + 1, rr);
    __syncthreads();
    {
        const Row4<T> rprev = L.get(sys, 2 * seg - 1);
        const Row4<T> lnext = L.get(sys, 2 * seg + 2);
        const Row4<T> nl = pcr_combine(rl, rprev, rr);
        const Row4<T> nr = pcr_combine(rr, rl, lnext);
        rl = nl;
        rr = nr;
    }
    // segment strides h = 1, 2, 4, ...: L rows and R rows are now separate systems
    for (int h = 1; h < stride_limit; h <<= 1) {
        __syncthreads();
        L.put(sys, 2 * seg, rl);
        L.put(sys, 2 * seg + 1, rr);
        __syncthreads();
        const Row4<T> llo = L.get(sys, 2 * (seg - h)), lhi = L.get(sys, 2 * (seg + h));
        const Row4<T> rlo = (seg - h >= 0) ? L.get(sys, 2 * (seg - h) + 1) : ident_row<T>();
        const Row4<T> rhi = (seg + h < P) ? L.get(sys, 2 * (seg + h) + 1) : ident_row<T>();
        rl = pcr_combine(rl, (seg - h >= 0) ? llo : ident_row<T>(),
                         (seg + h < P) ? lhi : ident_row<T>());
        rr = pcr_combine(rr, rlo, rhi);
    }
    __syncthreads();
}

// Thomas on the strided subsystems left by a PCR stopped at segment stride W:
// subsystem (sys, column, parity) = rows 2*(column + j*W) + parity, j < P/W.
// Called by every lane; `tid` < S*2*W lanes do one subsystem each.  Solutions
// are written to the d slot.
template <typename T>
__device__ __forceinline__ void strided_thomas(const IfaceLds<T> &L, int S, int P, int W, int tid) {
    const int per = 2 * W;
    if (tid < S * per) {
        const int sys = tid / per, col = (tid % per) >> 1, par = tid & 1;
        const int J = P / W;
        T cp = 0, dp = 0;
        for (int j = 0; j < J; ++j) {
            const int r = 2 * (col + j * W) + par;
            const int i = sys * L.N + r;
            const T a = L.a[i], b = L.b[i], c = L.c[i], d = L.d[i];
            const T k = rcp(b - a * cp);
            cp = c * k;
            dp = (d - a * dp) * k;
            L.c[i] = cp;
            L.d[i] = dp;
        }
        T x = 0;
        for (int j = J - 1; j >= 0; --j) {
            const int i = sys * L.N + 2 * (col + j * W) + par;
            x = L.d[i] - L.c[i] * x;
            L.d[i] = x;
        }
    }
    __syncthreads();
}

// Cyclic reduction (forward reduction to one row, then back substitution) on S
// systems of N = 2^k rows held in LDS, by all `lanes` lanes of the workgroup.
// x replaces d.
template <typename T>
__device__ __forceinline__ void cr_solve(const IfaceLds<T> &L, int S, int tid, int lanes) {
    const int N = L.N;
    // forward: rows i = 2h-1 (mod 2h) eliminate their neighbours at distance h
    for (int h = 1; h < N; h <<= 1) {
        const int act = N / (2 * h);
        for (int t = tid; t < S * act; t += lanes) {
            const int sys = t / act, i = (t % act) * 2 * h + 2 * h - 1;
            const Row4<T> x = L.get(sys, i);
            const Row4<T> y = pcr_combine(x, L.get(sys, i - h), L.get(sys, i + h));
            L.put(sys, i, y);
        }
        __syncthreads();
    }
    // back substitution: rows i = h-1 (mod 2h) from the solved rows at distance h
    for (int h = N; h >= 1; h >>= 1) {
        const int act = (h == N) ? 1 : N / (2 * h);
        for (int t = tid; t < S * act; t += lanes) {
            const int sys = t / act;
            const int i = (h == N) ? N - 1 : (t % act) * 2 * h + h - 1;
            const int base = sys * N;
            const T xm = (i - h >= 0 && h != N) ? L.d[base + i - h] : T(0);
            const T xp = (i + h < N && h != N) ? L.d[base + i + h] : T(0);
            L.d[base + i] = (L.d[base + i] - L.a[base + i] * xm - L.c[base + i] * xp) *
                            rcp(L.b[base + i]);
        }
        __syncthreads();
    }
}

}  // namespace tri
