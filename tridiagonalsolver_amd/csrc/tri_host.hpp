// Host-side plumbing shared by the kernel tables and the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include <string>

#include "tri_tile.hpp"

namespace tri {

enum Layout { kContiguous = 0, kPacked = 1, kInterleaved = 2 };

// One compiled tile kernel.
template <typename T> struct TileEntry {
    int M, NW, S;
    bool IL, VEC, ALDS, BIG, PAD, NT, XP;
    int cap;  // rows per system it can hold
    hipError_t (*launch)(const TileParams<T> &, hipStream_t);
};

// Largest LDS allocation of one workgroup on the current device.
int max_lds_per_block();

template <typename T, int M, int NW, int S, bool IL, bool VEC, bool ALDS, bool BIG, bool PAD,
          bool NT, bool XP>
hipError_t launch_tile(const TileParams<T> &p, hipStream_t st) {
    const int64_t tiles = (p.batch + S - 1) / S;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffff) return hipErrorInvalidValue;
    auto kern = tile_wm_kernel<T, M, NW, S, IL, VEC, ALDS, BIG, PAD, NT, XP>;
    // LDS for the interface solvers: 2 rows x 4 coefficients per lane (PCR/CR), or
    // 6 outer-form values per lane (Wang-Mou merges exchanged through LDS).
    size_t dyn = 0;
    if (iface_family(p.iface) >= kIfacePCR) dyn = (size_t)iface_lds_bytes<T>(64 * NW);
    else if (p.xchg_lds) dyn = (size_t)6 * 64 * NW * sizeof(T);
    if (dyn) {
        static size_t granted = 0;      // per instantiation
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
    TileParams<T> q = p;
    q.spread = 0;
    if constexpr (IL && S * (int)sizeof(T) >= 128) {
        // tiles of whole 128-byte row pieces, resident at once, spread over the batch
        // when rows start on lines and systems have at least 128 rows.  Measured on
        // MI355X (ms, in order / spread): FP32 256 x 65536 0.068 / 0.064, 512 x 32768
        // 0.078 / 0.072; FP64 0.136 / 0.129, 0.158 / 0.148; FP32 64 x 262144 0.082 /
        // 0.083.  Tiles of narrower pieces share lines with their neighbours
        // (xcd_group) and lost with it: FP64 2048 x 8191 0.227 / 0.269 ms.
        // ($TRIDIAG_TILE_SPREAD=0 turns it off.)
        constexpr int G = 1;
        static const int resident = [&] {
            int per_cu = 0, dev = 0, cus = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * NW, dyn) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
                return 0;
            return per_cu * cus;
        }();
        static const bool on = [] {
            const char *e = getenv("TRIDIAG_TILE_SPREAD");
            return !(e && atoi(e) == 0);
        }();
        if (on && p.n >= 128 && (p.rstride * (int64_t)sizeof(T)) % 128 == 0 && resident / G > 1)
            q.spread = resident / G;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)tiles), dim3(64 * NW), dyn, st, q);
    return hipGetLastError();
}

template <typename T> struct TileTable {
    const TileEntry<T> *entries;
    int count;
};

TileTable<float> tile_table_f32();
TileTable<double> tile_table_f64();

template <typename T> TileTable<T> tile_table();
template <> inline TileTable<float> tile_table<float>() { return tile_table_f32(); }
template <> inline TileTable<double> tile_table<double>() { return tile_table_f64(); }

// Systems of up to lane_thomas_max_rows() rows, one per lane, in registers.
int lane_thomas_max_rows(int precision);
hipError_t launch_lane_thomas_f32(const float *a, const float *b, const float *c, float *d,
                                  int64_t n, int64_t batch, int64_t sstride, int64_t rstride,
                                  bool il, int32_t bb, hipStream_t st);
hipError_t launch_lane_thomas_f64(const double *a, const double *b, const double *c, double *d,
                                  int64_t n, int64_t batch, int64_t sstride, int64_t rstride,
                                  bool il, int32_t bb, hipStream_t st);
// Layout conversion of up to four coefficient arrays: each source is a rows x cols
// matrix with row pitch ld_src, copied (transpose = false) or transposed into a
// destination with row pitch ld_dst.  col_tiles is filled by the launcher.
template <typename T> struct LayoutMove {
    const T *src[4];
    T *dst[4];
    int64_t rows, cols, ld_src, ld_dst, col_tiles;
};
hipError_t launch_layout_move_f32(const LayoutMove<float> &p, int count, bool transpose,
                                  hipStream_t st);
hipError_t launch_layout_move_f64(const LayoutMove<double> &p, int count, bool transpose,
                                  hipStream_t st);
// Thomas in the contiguous / packed layouts with 16-byte chunked rows per lane;
// hipErrorInvalidValue when rows are not 16-byte aligned or an array spans 2 GiB.
hipError_t launch_thomas_rows_f32(const float *a, const float *b, float *c, float *d, int64_t n,
                                  int64_t batch, int64_t sstride, int32_t bb, hipStream_t st);
hipError_t launch_thomas_rows_f64(const double *a, const double *b, double *c, double *d,
                                  int64_t n, int64_t batch, int64_t sstride, int32_t bb,
                                  hipStream_t st);
// Bulk-row forms of the reference's streaming solvers (tri_bulk.hip), one wave per
// system; form 0 PCR-pThomas, 1 PCR-half-pThomas, 2 WM-pGE, 3 WM-half-pGE; lds:
// lane exchanges through LDS instead of ds_bpermute.  Arrays below 2 GiB (bb =
// byte extent), sstride >= n.
hipError_t launch_bulk_f32(int form, bool lds, const float *a, const float *b, float *c, float *d,
                           int64_t n, int64_t batch, int64_t sstride, int32_t bb, hipStream_t st);
hipError_t launch_bulk_f64(int form, bool lds, const double *a, const double *b, double *c,
                           double *d, int64_t n, int64_t batch, int64_t sstride, int32_t bb,
                           hipStream_t st);
// The reference's CR id through global memory (tri_bulk.hip), one wave per system;
// scratch holds 4 * n * batch words (the reduced levels).
hipError_t launch_bulk_cr_f32(const float *a, const float *b, float *c, float *d, int64_t n,
                              int64_t batch, int64_t sstride, int32_t bb, float *scratch,
                              hipStream_t st);
hipError_t launch_bulk_cr_f64(const double *a, const double *b, double *c, double *d, int64_t n,
                              int64_t batch, int64_t sstride, int32_t bb, double *scratch,
                              hipStream_t st);
// The reference's register CR ids (CR_reg4/8/16, tri_bulk.hip): R rows per lane,
// one workgroup per system of at most 1024 R rows.
hipError_t launch_bulk_crreg_f32(int R, const float *a, const float *b, float *c, float *d,
                                 int64_t n, int64_t batch, int64_t sstride, int32_t bb,
                                 hipStream_t st);
hipError_t launch_bulk_crreg_f64(int R, const double *a, const double *b, double *c, double *d,
                                 int64_t n, int64_t batch, int64_t sstride, int32_t bb,
                                 hipStream_t st);
// Composed adjoint (systems the fused adjoint kernels do not hold): the shifted
// coefficients of A^T into ap / cp, g into lam, and d into xo when d is given;
// then, after the solves, the coefficient gradients from x and lambda.
hipError_t launch_adj_shift_f32(const float *a, const float *c, const float *g, const float *d,
                                float *ap, float *cp, float *lam, float *xo, int64_t n,
                                int64_t batch, bool il, hipStream_t st);
hipError_t launch_adj_shift_f64(const double *a, const double *c, const double *g,
                                const double *d, double *ap, double *cp, double *lam, double *xo,
                                int64_t n, int64_t batch, bool il, hipStream_t st);
hipError_t launch_adj_grad_f32(const float *x, const float *lam, float *ga, float *gb, float *gc,
                               int64_t n, int64_t batch, bool il, hipStream_t st);
hipError_t launch_adj_grad_f64(const double *x, const double *lam, double *ga, double *gb,
                               double *gc, int64_t n, int64_t batch, bool il, hipStream_t st);
hipError_t launch_thomas_f32(const float *a, const float *b, float *c, float *d, int64_t n,
                             int64_t batch, int64_t sstride, int64_t rstride, hipStream_t st);
hipError_t launch_thomas_f64(const double *a, const double *b, double *c, double *d, int64_t n,
                             int64_t batch, int64_t sstride, int64_t rstride, hipStream_t st);

}  // namespace tri
