// On-chip partition solver ("tile" kernels) for gfx950.
//
// A workgroup owns S whole systems.  Every system is cut into P segments of M
// consecutive rows; one lane owns one segment and keeps it in VGPRs from the
// moment it is loaded until its x is stored, so every unknown crosses HBM
// exactly 5 times (read a, b, c, d; write x).
//
//  1. lane sweep  : a Thomas-like forward sweep over the lane's rows writes each
//                   interior row as x_j = de_j - al_j x_L - ga_j x_{j+1}
//                   (x_L = the segment's first row) and leaves the segment's
//                   first and last rows coupled only to the rows just outside
//                   the segment (Wang's partition method, H.H. Wang 1981;
//                   the per-thread form of Laszlo, Giles, Appleyard, ACM TOMS
//                   42(4) 2016).
//  2. wave merge  : log2(segments per wave) Wang-Mou merge rounds (Wang & Mou
//                   1991) across lanes make the first/last row of every segment
//                   depend only on the rows just outside the wave-block.
//  3. WG chain    : a partial Gaussian elimination (pGE) chain over the
//                   wave-blocks of a system, in LDS, gives the x values at the
//                   wave-block edges.
//  4. substitution: every lane gets x_L, x_R of its segment and back-substitutes
//                   its interior rows from registers.
//
// The reference solves the same families with one warp streaming through a
// system (Tridiagonal_CUDA/tridiagonal_wm_pcr.cu:17-205 WM-pGE,
// :886-1018 PCR-pThomas), which writes c', d' to global memory and reads them
// back (9 words per unknown); this kernel keeps them on chip.
//
// Mappings (IL = interleaved layout, element (s, i) at i*batch + s):
//   IL : lane -> system lane % S, segment (wave, lane / S).  A wave reads
//        S consecutive systems per row: S*sizeof(T) = 32 contiguous bytes,
//        and the 128-byte line is shared by 128/(32) workgroups placed on the
//        same XCD (xcd_group below) so the line is read from HBM once.
//   !IL: contiguous per system (layout 0, or layout 1 with system stride 4n);
//        a system is owned by NWS consecutive waves, lane = segment.
#pragma once
#include "tri_device.hpp"

namespace tri {

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  Groups
// of G consecutive tile indices (which share 128-byte lines in the interleaved
// layout) are sent to one XCD in consecutive dispatch slots so that their L2
// holds the shared lines.
__device__ __forceinline__ int xcd_group(int b, int nblocks, int G) {
    const int span = 8 * G;
    const int nfull = (nblocks / span) * span;
    if (G <= 1 || b >= nfull) return b;
    const int xcd = b & 7, idx = b >> 3;
    const int q = idx / G, r = idx - q * G;
    return G * (8 * q + xcd) + r;
}

template <typename T> struct TileParams {
    const T *a, *b, *c;
    T *d;
    int64_t n, batch;
    int64_t sstride;  // distance between systems (IL: 1)
    int64_t rstride;  // distance between rows (IL: batch)
};

template <typename T, int M, int NW, int S, bool IL, bool VEC>
struct TileShape {
    static_assert(M >= 2, "a lane segment needs at least two rows");
    static constexpr int LS = IL ? S : 1;        // lane stride between segments of a system
    static constexpr int SW = 64 / LS;           // segments of one system per wave
    static constexpr int NWS = IL ? NW : NW / S; // waves per system
    static constexpr int P = SW * NWS;           // segments per system
    static constexpr int CAP = P * M;            // rows per system the tile can hold
    static constexpr int NI = (M > 2) ? (M - 2) : 1;
    static_assert(IL ? (64 % S == 0) : (NW % S == 0), "bad tile shape");
};

template <typename T, int M, int NW, int S, bool IL, bool VEC>
__global__ void __launch_bounds__(64 * NW) tile_wm_kernel(TileParams<T> prm) {
    using Sh = TileShape<T, M, NW, S, IL, VEC>;
    constexpr int SW = Sh::SW, NWS = Sh::NWS, LS = Sh::LS;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int sl = IL ? (lane % S) : (wave / NWS);   // system inside the workgroup
    const int kw = IL ? (lane / S) : lane;           // segment inside the wave-block
    const int ws = IL ? wave : (wave % NWS);         // wave-block inside the system
    const int64_t n = prm.n;
    int tile = blockIdx.x;
    if (IL) tile = xcd_group(blockIdx.x, gridDim.x, 128 / (S * (int)sizeof(T)));
    const int64_t sys = (int64_t)tile * S + sl;
    const bool valid = sys < prm.batch;
    const int64_t row0 = (int64_t)(ws * SW + kw) * M;
    const int64_t base = valid ? sys * prm.sstride : 0;
    const int64_t rs = prm.rstride;
    const T *__restrict__ A = prm.a;
    const T *__restrict__ B = prm.b;
    const T *__restrict__ C = prm.c;
    T *__restrict__ D = prm.d;

    // ---------------------------------------------------------------- 1. sweep
    T al[Sh::NI], ga[Sh::NI], de[Sh::NI];
    T aL = 0, bL = 1, cL = 0, dL = 0, aR = 0, bR = 1, cR = 0, dR = 0;
    T g = 0, alp = -1, del = 0;  // virtual row before the first interior row

    constexpr int V = (VEC && !IL) ? Vec16<T>::N : 1;
    static_assert(M % V == 0, "vector width must divide M");
#pragma unroll
    for (int j0 = 0; j0 < M; j0 += V) {
        T va[V], vb[V], vc[V], vd[V];
        const int64_t r0 = row0 + j0;
        if (V > 1 && valid && r0 + V <= n) {
            using VT = typename Vec16<T>::type;
            const int64_t off = base + r0;
            VT xa = *reinterpret_cast<const VT *>(A + off);
            VT xb = *reinterpret_cast<const VT *>(B + off);
            VT xc = *reinterpret_cast<const VT *>(C + off);
            VT xd = *reinterpret_cast<const VT *>(D + off);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                va[v] = reinterpret_cast<const T *>(&xa)[v];
                vb[v] = reinterpret_cast<const T *>(&xb)[v];
                vc[v] = reinterpret_cast<const T *>(&xc)[v];
                vd[v] = reinterpret_cast<const T *>(&xd)[v];
            }
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const int64_t r = r0 + v;
                if (valid && r < n) {
                    const int64_t off = base + r * rs;
                    va[v] = A[off];
                    vb[v] = B[off];
                    vc[v] = C[off];
                    vd[v] = D[off];
                } else {
                    va[v] = 0; vb[v] = 1; vc[v] = 0; vd[v] = 0;
                }
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int j = j0 + v;
            const int64_t r = r0 + v;
            T a = (r == 0) ? T(0) : va[v];          // a_0 is not part of A
            T c = (r == n - 1) ? T(0) : vc[v];      // nor is c_{n-1}
            T b = vb[v], d = vd[v];
            if (j == 0) {
                aL = a; bL = b; cL = c; dL = d;
            } else if (j == M - 1) {
                aR = a; bR = b; cR = c; dR = d;
            } else {
                T rr = rcp(b - a * g);
                g = c * rr;
                alp = -(a * alp) * rr;
                del = (d - a * del) * rr;
                al[j - 1] = alp;
                ga[j - 1] = g;
                de[j - 1] = del;
            }
        }
    }
    // Row L from x_1 = DD - AA x_L - GG x_R (backward over the interior).
    T DD = 0, AA = 0, GG = -1;
    if constexpr (M > 2) {
        DD = de[M - 3]; AA = al[M - 3]; GG = ga[M - 3];
#pragma unroll
        for (int j = M - 4; j >= 0; --j) {
            DD = de[j] - ga[j] * DD;
            AA = al[j] - ga[j] * AA;
            GG = -ga[j] * GG;
        }
    }
    // Reduced rows of the segment:
    //   L: aL u + DLr x_L + ELr x_R = FLr ;  R: ARr x_L + DRr x_R + cR v = FRr
    const T DLr = bL - cL * AA, ELr = -(cL * GG), FLr = dL - cL * DD;
    const T ARr = -(aR * alp), DRr = bR - aR * g, FRr = dR - aR * del;
    Outer<T> oL, oR;
    {
        T inv = rcp(DLr * DRr - ELr * ARr);
        oL.r = (DRr * FLr - ELr * FRr) * inv;
        oL.p = (DRr * aL) * inv;
        oL.q = -(ELr * cR) * inv;
        oR.r = (DLr * FRr - ARr * FLr) * inv;
        oR.p = -(ARr * aL) * inv;
        oR.q = (DLr * cR) * inv;
    }

    // ---------------------------------------------------------------- 2. wave merge
#pragma unroll
    for (int t = 0; (1 << t) < SW; ++t) {
        const int h = 1 << t;
        const int blk = kw >> t;
        const int pb = (blk & ~1) << t;              // first segment of the merged pair
        const int srcA = sl % LS + LS * (pb + h - 1); // lane with A's last row
        const int srcB = sl % LS + LS * (pb + h);     // lane with B's first row
        T rA = lane_read(oR.r, srcA), pA = lane_read(oR.p, srcA), qA = lane_read(oR.q, srcA);
        T rB = lane_read(oL.r, srcB), pB = lane_read(oL.p, srcB), qB = lane_read(oL.q, srcB);
        T RB, PB, QB, RA, PA, QA;
        wm_pair_coeffs(rA, pA, qA, rB, pB, qB, RB, PB, QB, RA, PA, QA);
        const bool inA = (blk & 1) == 0;
        wm_merge_row(oL, inA, RB, PB, QB, RA, PA, QA);
        wm_merge_row(oR, inA, RB, PB, QB, RA, PA, QA);
    }

    // ---------------------------------------------------------------- 3. WG chain
    T u = 0, v = 0;
    if constexpr (NWS > 1) {
        __shared__ T sh_blk[S][NWS][6];
        __shared__ T sh_uv[S][NWS][2];
        if (kw == 0) {
            sh_blk[sl][ws][0] = oL.r; sh_blk[sl][ws][1] = oL.p; sh_blk[sl][ws][2] = oL.q;
        }
        if (kw == SW - 1) {
            sh_blk[sl][ws][3] = oR.r; sh_blk[sl][ws][4] = oR.p; sh_blk[sl][ws][5] = oR.q;
        }
        __syncthreads();
        if (threadIdx.x < S) {
            const int q = threadIdx.x;
            T rho = 0, sig = 0;
            // forward: x_{E_{w-1}} = rho - sig x_{F_w}
#pragma unroll
            for (int w = 0; w < NWS; ++w) {
                T rF = sh_blk[q][w][0], pF = sh_blk[q][w][1], qF = sh_blk[q][w][2];
                T rE = sh_blk[q][w][3], pE = sh_blk[q][w][4], qE = sh_blk[q][w][5];
                T inv = rcp(T(1) - pF * sig);
                T rhoF = (rF - pF * rho) * inv;
                T sigF = qF * inv;
                T rhoE = rE - pE * (rho - sig * rhoF);
                T sigE = qE + pE * sig * sigF;
                sh_blk[q][w][0] = rhoF; sh_blk[q][w][1] = sigF;
                sh_blk[q][w][2] = rhoE; sh_blk[q][w][3] = sigE;
                rho = rhoE; sig = sigE;
            }
            // backward: x_{F_w} and x_{E_w} from x_{F_{w+1}}
            T xFn = 0;
            sh_uv[q][0][0] = 0;
#pragma unroll
            for (int w = NWS - 1; w >= 0; --w) {
                T xE = sh_blk[q][w][2] - sh_blk[q][w][3] * xFn;
                T xF = sh_blk[q][w][0] - sh_blk[q][w][1] * xFn;
                sh_uv[q][w][1] = xFn;
                if (w + 1 < NWS) sh_uv[q][w + 1][0] = xE;
                xFn = xF;
            }
        }
        __syncthreads();
        u = sh_uv[sl][ws][0];
        v = sh_uv[sl][ws][1];
    }

    // ---------------------------------------------------------------- 4. substitution
    const T xL = oL.r - oL.p * u - oL.q * v;
    const T xR = oR.r - oR.p * u - oR.q * v;
    if constexpr (M > 2) {
        T xn = xR;
#pragma unroll
        for (int j = M - 3; j >= 0; --j) {
            xn = de[j] - al[j] * xL - ga[j] * xn;
            de[j] = xn;
        }
    }
    if (!valid) return;
#pragma unroll
    for (int j0 = 0; j0 < M; j0 += V) {
        T xv[V];
#pragma unroll
        for (int vv = 0; vv < V; ++vv) {
            const int j = j0 + vv;
            xv[vv] = (j == 0) ? xL : ((j == M - 1) ? xR : de[(j >= 1 && j <= M - 2) ? j - 1 : 0]);
        }
        const int64_t r0 = row0 + j0;
        if (V > 1 && r0 + V <= n) {
            using VT = typename Vec16<T>::type;
            VT xs;
#pragma unroll
            for (int vv = 0; vv < V; ++vv) reinterpret_cast<T *>(&xs)[vv] = xv[vv];
            *reinterpret_cast<VT *>(D + base + r0) = xs;
        } else {
#pragma unroll
            for (int vv = 0; vv < V; ++vv) {
                const int64_t r = r0 + vv;
                if (r < n) D[base + r * rs] = xv[vv];
            }
        }
    }
}

// Naive batched Thomas: one lane per system, the recurrence of
// Tridiagonal_CUDA/tridiagonal_basic.cu:4-30.  Writes c' into c and x into d
// (the reference's in-place semantics).  Coalesced in the interleaved layout,
// lane-strided (one cache line per lane and row) in the contiguous layout.
template <typename T>
__global__ void __launch_bounds__(64) thomas_kernel(const T *__restrict__ A,
                                                    const T *__restrict__ B, T *__restrict__ C,
                                                    T *__restrict__ D, int64_t n, int64_t batch,
                                                    int64_t sstride, int64_t rs) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= batch) return;
    const int64_t base = s * sstride;
    T rb = rcp(B[base]);
    T cc = (n > 1) ? C[base] * rb : T(0);
    T dd = D[base] * rb;
    C[base] = cc;
    D[base] = dd;
    for (int64_t i = 1; i < n; ++i) {
        const int64_t o = base + i * rs;
        T a = A[o];
        T k = rcp(B[o] - cc * a);
        cc = (i < n - 1) ? C[o] * k : T(0);
        dd = (D[o] - dd * a) * k;
        C[o] = cc;
        D[o] = dd;
    }
    for (int64_t i = n - 2; i >= 0; --i) {
        const int64_t o = base + i * rs;
        dd = D[o] - C[o] * dd;
        D[o] = dd;
    }
}

}  // namespace tri
