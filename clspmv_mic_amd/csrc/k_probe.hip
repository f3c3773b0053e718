// HBM bandwidth probe: a streaming float4 copy (replaces the reference's
// gpu_memcpy probe, kernels/mem_bandwidth.cl, src/mem_bandwidth.cpp).  Used
// to calibrate the roofline and the hardware byte counters.

#include "device_common.hpp"

namespace clspmv {
namespace {

using namespace dev;

__global__ __launch_bounds__(kBlock) void copy4_kernel(long long n4, const float4* __restrict__ src,
                                                       float4* __restrict__ dst) {
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n4; i += stride)
        dst[i] = ld_stream(src + i);
}

// y[i] = x[idx(i)] with idx a hash of i: the random-gather ceiling of a
// kernel whose x reads have no locality.  Streams 4 bytes out per gather.
__global__ __launch_bounds__(kBlock) void gather_kernel(long long n, int nx, const float* __restrict__ x,
                                                        float* __restrict__ y) {
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        unsigned long long v = (unsigned long long)i * 0x9E3779B97F4A7C15ull;
        v ^= v >> 29;
        y[i] = x[(unsigned)(v % (unsigned)nx)];
    }
}

}  // namespace

namespace launch {

hipError_t gather(hipStream_t s, long long n, int nx, const float* x, float* y) {
    if (n <= 0 || nx <= 0) return hipSuccess;
    long long want = (n + kBlock - 1) / kBlock;
    int grid = (int)(want < 256 * 64 ? want : 256 * 64);
    hipLaunchKernelGGL(gather_kernel, dim3(grid), dim3(kBlock), 0, s, n, nx, x, y);
    return hipGetLastError();
}

hipError_t copy4(hipStream_t s, long long n4, const float* src, float* dst) {
    if (n4 <= 0) return hipSuccess;
    long long want = (n4 + kBlock - 1) / kBlock;
    int grid = (int)(want < 256 * 64 ? want : 256 * 64);
    hipLaunchKernelGGL(copy4_kernel, dim3(grid), dim3(kBlock), 0, s, n4,
                       reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst));
    return hipGetLastError();
}

}  // namespace launch
}  // namespace clspmv
