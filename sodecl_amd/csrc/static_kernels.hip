// Kernels compiled ahead of time for gfx950 (everything that depends on the
// user's system is compiled at run time by hipRTC instead).
#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/sodecl.h"

// Plain streaming copy, 16 bytes per lane per iteration, grid-stride: the
// reference point for reading HBM traffic counters (bytes fetched and written
// by a kernel of known size).
__global__ void __launch_bounds__(256) sodecl_calibration_copy(const double2* __restrict__ src,
                                                                double2* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

extern "C" int sodecl_copy_device(void* dst, const void* src, size_t bytes, void* stream)
{
    if (!dst || !src || bytes % 16 != 0) return 0;
    const size_t n = bytes / 16;
    if (n == 0) return 1;
    int device = 0;
    (void)hipGetDevice(&device);
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    size_t blocks = (n + 255) / 256;
    const size_t cap = (size_t)cus * 8;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(sodecl_calibration_copy, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const double2*)src, (double2*)dst, n);
    return hipGetLastError() == hipSuccess ? 1 : 0;
}
