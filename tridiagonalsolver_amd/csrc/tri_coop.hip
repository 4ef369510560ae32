// Cooperative row-block kernels: instantiations, launchers and their table.
#include <algorithm>
#include <mutex>

#include "tri_coop.hpp"

namespace tri {

template <typename T, int M, int NW, int S, bool IL, int KMAX, bool PF, int OCC>
static hipError_t launch_coop(const CoopParams<T> &p, int grid, hipStream_t st) {
    if (grid < 1 || p.K < 1 || p.K > KMAX || grid % p.K != 0) return hipErrorInvalidValue;
    if (p.pipelined)
        hipLaunchKernelGGL((coop_pipe_kernel<T, M, NW, S, IL, KMAX>), dim3((unsigned)grid),
                           dim3(64 * NW), 0, st, p);
    else
        hipLaunchKernelGGL((coop_kernel<T, M, NW, S, IL, KMAX, PF, OCC>), dim3((unsigned)grid),
                           dim3(64 * NW), 0, st, p);
    return hipGetLastError();
}

// Workgroups of the entry's kernel (pipelined or not) that the current device holds
// at once (0 if unknown), cached per device.
template <typename T, int M, int NW, int S, bool IL, int KMAX, bool PF, int OCC>
static int coop_max_grid(bool pipelined) {
    constexpr int kDevs = 64;
    static std::mutex mu;
    static int cache[kDevs][2] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevs) return 0;
    std::lock_guard<std::mutex> lk(mu);
    int &c = cache[dev][pipelined ? 1 : 0];
    if (c == 0) {
        int per_cu = 0, cus = 0;
        const void *fn = pipelined
            ? reinterpret_cast<const void *>(coop_pipe_kernel<T, M, NW, S, IL, KMAX>)
            : reinterpret_cast<const void *>(coop_kernel<T, M, NW, S, IL, KMAX, PF, OCC>);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * NW, 0) != hipSuccess)
            return 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        c = per_cu * cus;
    }
    return c;
}

#define TRI_COOP_ENTRY(T, M, NW, S, IL, KMAX, PIPE, PF, OCC)                                 \
    {M, NW, S, IL, CoopShape<T, M, NW, S, IL>::R, KMAX, PIPE,                                 \
     &launch_coop<T, M, NW, S, IL, KMAX, PF, OCC>, &coop_max_grid<T, M, NW, S, IL, KMAX, PF, OCC>}

// Interleaved: S systems = one 128-byte line per row, blocks of 2*NW*M (FP32) or
// 4*NW*M (FP64) rows, up to 32 blocks, and an entry for 33-64 blocks whose larger
// block-form staging leaves one workgroup per CU (interleaved 12288 / 16384 rows x
// 1024: FP32 0.129 / 0.167 ms, FP64 0.266 / 0.367 ms, against 0.164 / 0.213 and
// 0.363 / 0.477 ms through a contiguous copy).  Contiguous: one system, blocks of
// 64*NW*M rows, up to 32 blocks (a group of 64 measured slower than the chunked
// solve).  PIPE: coop_pipe_kernel by default; PF: the
// next block in flight in registers; OCC: waves per SIMD.
// Measured on MI355X (ms; 16.7 M unknowns unless noted):
//   interleaved  FP32 4096 x 4096: PF, 1 workgroup per CU 0.117; pipelined 0.109;
//                no PF, 2 per CU (OCC 4) 0.108.  FP32 2048 x 8192: 0.102 / 0.091
//                (pipelined / OCC 4); 8192 x 2048: 0.129 / 0.123.  FP64 4096 x
//                4096: 0.233 / 0.223 / 0.214 (PF / pipelined / OCC 4).  Blocks of
//                128 rows (M = 8 FP32, 4 FP64) with K = 32: 0.148 / 0.282; 16-wave
//                workgroups of M = 8 / 4 (no spills, one per CU): 0.133 / 0.286.
//   contiguous   FP32 20000 rows: PF OCC 2 0.083, OCC 4 0.082, OCC 3 0.076;
//                131072 rows (32 blocks): OCC 4 0.093, OCC 3 0.088, chunked 0.107.
//                FP64 20000 rows: PF OCC 2 0.180, OCC 4 0.147, OCC 3 0.151, chunked
//                0.209; 65536 x 256 (32 blocks): OCC 4 0.173, chunked 0.206.
// The register prefetch costs more occupancy than it hides latency; more
// workgroups per CU keep more loads in flight, even where the kernel spills.
static const CoopEntry<float> kCoopF32[] = {
    TRI_COOP_ENTRY(float, 16, 8, 32, true, 32, false, false, 4),
    TRI_COOP_ENTRY(float, 16, 8, 32, true, 64, false, false, 4),
    TRI_COOP_ENTRY(float, 16, 4, 1, false, 32, false, false, 3),
};
static const CoopEntry<double> kCoopF64[] = {
    TRI_COOP_ENTRY(double, 8, 8, 16, true, 32, false, false, 4),
    TRI_COOP_ENTRY(double, 8, 8, 16, true, 64, false, false, 4),
    TRI_COOP_ENTRY(double, 8, 4, 1, false, 32, false, false, 4),
};
#undef TRI_COOP_ENTRY

CoopTable<float> coop_table_f32() { return {kCoopF32, (int)(sizeof(kCoopF32) / sizeof(kCoopF32[0]))}; }
CoopTable<double> coop_table_f64() {
    return {kCoopF64, (int)(sizeof(kCoopF64) / sizeof(kCoopF64[0]))};
}

}  // namespace tri
