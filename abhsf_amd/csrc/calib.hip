// calib.hip -- plain device copy kernel of known size: the HBM calibration
// point for roofline and rocprofv3 FETCH_SIZE / WRITE_SIZE scaling (the
// STREAM-copy pattern of the reference's host probe mb/mb.cpp:31-48).
#include <hip/hip_runtime.h>

#include <cstddef>

__global__ __launch_bounds__(256) void k_stream_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                     size_t n2)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n2; i += stride) dst[i] = src[i];
}

extern "C" int abhsf_calib_copy(const double *src, double *dst, size_t n, void *stream)
{
    if (n % 2) return 1;
    const size_t n2 = n / 2;
    unsigned grid = (unsigned)((n2 + 255) / 256);
    if (grid > 256 * 8 * 16) grid = 256 * 8 * 16;
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const double2 *>(src), reinterpret_cast<double2 *>(dst), n2);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
