// Host-side plumbing shared by the kernel tables and the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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

template <typename T, int M, int NW, int S, bool IL, bool VEC, bool ALDS, bool BIG, bool PAD,
          bool NT, bool XP>
hipError_t launch_tile(const TileParams<T> &p, hipStream_t st) {
    const int64_t tiles = (p.batch + S - 1) / S;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((tile_wm_kernel<T, M, NW, S, IL, VEC, ALDS, BIG, PAD, NT, XP>),
                       dim3((unsigned)tiles),
                       dim3(64 * NW), 0, st, p);
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

hipError_t launch_thomas_f32(const float *a, const float *b, float *c, float *d, int64_t n,
                             int64_t batch, int64_t sstride, int64_t rstride, hipStream_t st);
hipError_t launch_thomas_f64(const double *a, const double *b, double *c, double *d, int64_t n,
                             int64_t batch, int64_t sstride, int64_t rstride, hipStream_t st);

}  // namespace tri
