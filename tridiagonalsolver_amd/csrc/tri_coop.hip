// Cooperative row-block kernels: instantiations, launchers and their table.
#include <algorithm>
#include <mutex>

#include "tri_coop.hpp"

namespace tri {

template <typename T, int M, int NW, int S, bool IL, int KMAX>
static hipError_t launch_coop(const CoopParams<T> &p, int grid, hipStream_t st) {
    if (grid < 1 || p.K < 1 || p.K > KMAX || grid % p.K != 0) return hipErrorInvalidValue;
    if (p.pipelined)
        hipLaunchKernelGGL((coop_pipe_kernel<T, M, NW, S, IL, KMAX>), dim3((unsigned)grid),
                           dim3(64 * NW), 0, st, p);
    else
        hipLaunchKernelGGL((coop_kernel<T, M, NW, S, IL, KMAX>), dim3((unsigned)grid),
                           dim3(64 * NW), 0, st, p);
    return hipGetLastError();
}

// Workgroups of the kernel that the current device holds at once (0 if unknown),
// cached per device.
template <typename T, int M, int NW, int S, bool IL, int KMAX> static int coop_max_grid() {
    constexpr int kDevs = 64;
    static std::mutex mu;
    static int cache[kDevs] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevs) return 0;
    std::lock_guard<std::mutex> lk(mu);
    if (cache[dev] == 0) {
        int per_cu = 0, cus = 0;
        // both kernels of the entry must fit the grid
        int per_cu_pipe = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void *>(coop_kernel<T, M, NW, S, IL, KMAX>), 64 * NW,
                0) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu_pipe,
                reinterpret_cast<const void *>(coop_pipe_kernel<T, M, NW, S, IL, KMAX>), 64 * NW,
                0) != hipSuccess)
            return 0;
        per_cu = std::min(per_cu, per_cu_pipe);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        cache[dev] = per_cu * cus;
    }
    return cache[dev];
}

#define TRI_COOP_ENTRY(T, M, NW, S, IL, KMAX, PIPE)                                          \
    {M, NW, S, IL, CoopShape<T, M, NW, S, IL>::R, KMAX, PIPE,                                 \
     &launch_coop<T, M, NW, S, IL, KMAX>, &coop_max_grid<T, M, NW, S, IL, KMAX>}

// Interleaved: S systems = one 128-byte line per row, blocks of 2*NW*M (FP32) or
// 4*NW*M (FP64) rows, up to 32 blocks.  Contiguous: one system, blocks of
// 64*NW*M rows, up to 16 blocks.  Measured on MI355X at 2^24 unknowns these were
// the fastest shapes; 4-wave interleaved workgroups (2 per CU, 32 blocks of 128
// rows at n = 4096) and 8-wave contiguous ones were slower at every size.
// PIPE: the entry runs coop_pipe_kernel (waits one task later).  Measured (ms,
// pipelined / waiting): FP64 interleaved 4096 x 4096 0.226 / 0.233; FP32
// interleaved 4096 x 4096 0.111 / 0.113, 8192 x 2048 0.129 / 0.132; contiguous
// FP32 20000 rows 0.098 / 0.083, FP64 0.190 / 0.178, FP32 65536 rows 0.100 / 0.086
// (the pipelined contiguous kernels spill).
static const CoopEntry<float> kCoopF32[] = {
    TRI_COOP_ENTRY(float, 16, 8, 32, true, 32, true),
    TRI_COOP_ENTRY(float, 16, 4, 1, false, 16, false),
};
static const CoopEntry<double> kCoopF64[] = {
    TRI_COOP_ENTRY(double, 8, 8, 16, true, 32, true),
    TRI_COOP_ENTRY(double, 8, 4, 1, false, 16, false),
};
#undef TRI_COOP_ENTRY

CoopTable<float> coop_table_f32() { return {kCoopF32, (int)(sizeof(kCoopF32) / sizeof(kCoopF32[0]))}; }
CoopTable<double> coop_table_f64() {
    return {kCoopF64, (int)(sizeof(kCoopF64) / sizeof(kCoopF64[0]))};
}

}  // namespace tri
