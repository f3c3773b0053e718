// HBM bandwidth probe: a streaming float4 copy (replaces the reference's
// gpu_memcpy probe, kernels/mem_bandwidth.cl, src/mem_bandwidth.cpp).  Used
// to calibrate the roofline and the hardware byte counters.

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;

__global__ __launch_bounds__(kBlock) void copy4_kernel(long long n4, const float4* __restrict__ src,
                                                       float4* __restrict__ dst) {
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n4; i += stride)
        dst[i] = ld_stream(src + i);
}

// The same copy with U float4 per thread per step, all U loads issued
// before the first store, and non-temporal stores: more bytes in flight
// per wave, for the launch shapes the probe sweeps.
template <int U>
__global__ __launch_bounds__(kBlock) void copy4u_kernel(long long n4, const float4* __restrict__ src,
                                                        float4* __restrict__ dst) {
    const long long stride = (long long)gridDim.x * kBlock * U;
    for (long long b = (long long)blockIdx.x * kBlock * U + threadIdx.x; b < n4; b += stride) {
        f32x4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = min(b + (long long)u * kBlock, n4 - 1);
            v[u] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(src + i));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = b + (long long)u * kBlock;
            if (i < n4) __builtin_nontemporal_store(v[u], reinterpret_cast<f32x4*>(dst + i));
        }
    }
}

// Read-only stream: every float4 of src is loaded once (U per thread per
// step); each thread writes one sum so the loads are not dropped.  The
// ceiling of kernels that mostly read, as SpMV does.
template <int U>
__global__ __launch_bounds__(kBlock) void read4u_kernel(long long n4, const float4* __restrict__ src,
                                                        float* __restrict__ out) {
    const long long stride = (long long)gridDim.x * kBlock * U;
    float acc = 0.0f;
    for (long long b = (long long)blockIdx.x * kBlock * U + threadIdx.x; b < n4; b += stride) {
        f32x4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = min(b + (long long)u * kBlock, n4 - 1);
            v[u] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(src + i));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    }
    out[(long long)blockIdx.x * kBlock + threadIdx.x] = acc;
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

hipError_t copy4u(hipStream_t s, long long n4, const float* src, float* dst, int unroll, int grid) {
    if (n4 <= 0) return hipSuccess;
    const auto* a = reinterpret_cast<const float4*>(src);
    auto* b = reinterpret_cast<float4*>(dst);
    switch (unroll) {
        case 2: hipLaunchKernelGGL(copy4u_kernel<2>, dim3(grid), dim3(kBlock), 0, s, n4, a, b); break;
        case 4: hipLaunchKernelGGL(copy4u_kernel<4>, dim3(grid), dim3(kBlock), 0, s, n4, a, b); break;
        case 8: hipLaunchKernelGGL(copy4u_kernel<8>, dim3(grid), dim3(kBlock), 0, s, n4, a, b); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t read4u(hipStream_t s, long long n4, const float* src, float* out, int unroll, int grid) {
    if (n4 <= 0) return hipSuccess;
    const auto* a = reinterpret_cast<const float4*>(src);
    switch (unroll) {
        case 2: hipLaunchKernelGGL(read4u_kernel<2>, dim3(grid), dim3(kBlock), 0, s, n4, a, out); break;
        case 4: hipLaunchKernelGGL(read4u_kernel<4>, dim3(grid), dim3(kBlock), 0, s, n4, a, out); break;
        case 8: hipLaunchKernelGGL(read4u_kernel<8>, dim3(grid), dim3(kBlock), 0, s, n4, a, out); break;
        default: return hipErrorInvalidValue;
    }
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
