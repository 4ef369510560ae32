// Large-system (chunked, three-pass) kernels: instantiations and launchers.
#include "tri_large.hpp"

namespace tri {

template <typename T, int M, int NW>
static hipError_t launch_large_shape(const LargeParams<T> &p, hipStream_t st) {
    if (p.K < 1 || p.K > 1024 || p.batch < 1 || p.batch > 65535) return hipErrorInvalidValue;
    size_t dyn = 0;
    if (iface_family(p.iface) >= kIfacePCR) dyn = (size_t)iface_lds_bytes<T>(64 * NW);
    auto solve = chunk_kernel<T, M, NW, kChunkSolve>;
    if (dyn) {
        static size_t granted = 0;
        if (dyn > granted) {
            hipFuncAttributes fa;
            hipError_t e = hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(solve));
            if (e != hipSuccess) return e;
            if (fa.sharedSizeBytes + dyn > (size_t)max_lds_per_block())
                return hipErrorInvalidConfiguration;
            e = hipFuncSetAttribute(reinterpret_cast<const void *>(solve),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
            if (e != hipSuccess) return e;
            granted = dyn;
        }
    }
    const dim3 grid((unsigned)p.K, (unsigned)p.batch);
    hipLaunchKernelGGL((chunk_kernel<T, M, NW, kChunkReduce>), grid, dim3(64 * NW), 0, st, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int lanes = 64 * ((p.K + 63) / 64);
    hipLaunchKernelGGL(chunk_iface_kernel<T>, dim3((unsigned)p.batch), dim3(lanes), 0, st, p);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(solve, grid, dim3(64 * NW), dyn, st, p);
    return hipGetLastError();
}

// Chunk shapes {M, NW}: chunk rows C = 64*NW*M.
static const LargeEntry<float> kLargeF32[] = {
    {16, 8, 8192, &launch_large_shape<float, 16, 8>},
    {8, 8, 4096, &launch_large_shape<float, 8, 8>},
};
static const LargeEntry<double> kLargeF64[] = {
    {8, 8, 4096, &launch_large_shape<double, 8, 8>},
    {4, 8, 2048, &launch_large_shape<double, 4, 8>},
};

int large_shape_count(int precision) {
    return precision == 0 ? (int)(sizeof(kLargeF32) / sizeof(kLargeF32[0]))
                          : (int)(sizeof(kLargeF64) / sizeof(kLargeF64[0]));
}
int large_chunk_rows(int precision, int shape) {
    if (shape < 0 || shape >= large_shape_count(precision)) return 0;
    return precision == 0 ? kLargeF32[shape].C : kLargeF64[shape].C;
}
hipError_t launch_large_f32(const LargeParams<float> &p, int shape, hipStream_t st) {
    if (shape < 0 || shape >= large_shape_count(0)) return hipErrorInvalidValue;
    return kLargeF32[shape].launch(p, st);
}
hipError_t launch_large_f64(const LargeParams<double> &p, int shape, hipStream_t st) {
    if (shape < 0 || shape >= large_shape_count(1)) return hipErrorInvalidValue;
    return kLargeF64[shape].launch(p, st);
}

}  // namespace tri
