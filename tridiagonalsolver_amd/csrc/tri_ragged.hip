// Ragged persistent kernel: instantiations and launchers.
#include "tri_ragged.hpp"

namespace tri {

// Persistent grid: as many workgroups as fit on the device at once.
template <typename T> static int ragged_grid() {
    static int g = [] {
        int dev = 0, cus = 256, per = 1;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per, reinterpret_cast<const void *>(ragged_kernel<T>), kRaggedLanes,
                ragged_lds_bytes<T>()) != hipSuccess ||
            per < 1)
            per = 1;
        return cus * per;
    }();
    return g;
}
int ragged_grid_f32() { return ragged_grid<float>(); }
int ragged_grid_f64() { return ragged_grid<double>(); }

template <typename T>
static hipError_t launch_ragged(const RaggedParams<T> &p, int grid, hipStream_t st) {
    static const hipError_t granted = [] {
        return hipFuncSetAttribute(reinterpret_cast<const void *>(ragged_kernel<T>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   ragged_lds_bytes<T>());
    }();
    if (granted != hipSuccess) return granted;
    hipLaunchKernelGGL(ragged_kernel<T>, dim3(grid), dim3(kRaggedLanes), ragged_lds_bytes<T>(), st,
                       p);
    return hipGetLastError();
}
hipError_t launch_ragged_f32(const RaggedParams<float> &p, int grid, hipStream_t st) {
    return launch_ragged<float>(p, grid, st);
}
hipError_t launch_ragged_f64(const RaggedParams<double> &p, int grid, hipStream_t st) {
    return launch_ragged<double>(p, grid, st);
}

}  // namespace tri
