// ELL and SELL SpMV for gfx950 (replace kernels/spmv_ell.cl and
// kernels/spmv_sell.cl).
//
// Both formats are column-major inside a slab (ELL: one slab of
// `ell_height_aligned` rows; SELL: one slab per slice), so consecutive lanes
// own consecutive rows and every slot load is one contiguous 256-byte wave
// access.  The 4-rows-per-lane ELL variant issues 128-bit loads.

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;

__global__ __launch_bounds__(kBlock) void ell_kernel(
    int rows, int hal, int ellnum, const int* __restrict__ col,
    const float* __restrict__ val, const float* __restrict__ x,
    float* __restrict__ y) {
    const int r = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (r >= rows) return;
    float a0 = 0.0f, a1 = 0.0f;
    int s = 0;
    for (; s + 1 < ellnum; s += 2) {
        const size_t i0 = (size_t)s * hal + r, i1 = i0 + hal;
        const int c0 = ld_stream(col + i0), c1 = ld_stream(col + i1);
        a0 = fmaf(ld_stream(val + i0), x[c0], a0);
        a1 = fmaf(ld_stream(val + i1), x[c1], a1);
    }
    if (s < ellnum) {
        const size_t i0 = (size_t)s * hal + r;
        a0 = fmaf(ld_stream(val + i0), x[ld_stream(col + i0)], a0);
    }
    y[r] += a0 + a1;
}

// Four consecutive rows per lane; requires hal % 4 == 0.
__global__ __launch_bounds__(kBlock) void ell4_kernel(
    int rows, int hal, int ellnum, const int* __restrict__ col,
    const float* __restrict__ val, const float* __restrict__ x,
    float* __restrict__ y) {
    const int r = (xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x) * 4;
    if (r >= rows) return;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int s = 0; s < ellnum; ++s) {
        const size_t i = (size_t)s * hal + r;
        const int4 c = ld_stream(reinterpret_cast<const int4*>(col + i));
        const float4 v = ld_stream(reinterpret_cast<const float4*>(val + i));
        a.x = fmaf(v.x, x[c.x], a.x);
        a.y = fmaf(v.y, x[c.y], a.y);
        a.z = fmaf(v.z, x[c.z], a.z);
        a.w = fmaf(v.w, x[c.w], a.w);
    }
    y[r] += a.x;
    if (r + 1 < rows) y[r + 1] += a.y;
    if (r + 2 < rows) y[r + 2] += a.z;
    if (r + 3 < rows) y[r + 3] += a.w;
}

__global__ __launch_bounds__(kBlock) void sell_kernel(
    int rows, int h, const int* __restrict__ ptr, const int* __restrict__ col,
    const float* __restrict__ val, const float* __restrict__ x,
    float* __restrict__ y) {
    const int r = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (r >= rows) return;
    const int sl = r / h, i = r - sl * h;
    const int p0 = ptr[sl];
    const int w = (ptr[sl + 1] - p0) / h;
    const int* c = col + p0 + i;
    const float* v = val + p0 + i;
    float a0 = 0.0f, a1 = 0.0f;
    int k = 0;
    for (; k + 1 < w; k += 2) {
        const int c0 = ld_stream(c + (size_t)k * h), c1 = ld_stream(c + (size_t)(k + 1) * h);
        a0 = fmaf(ld_stream(v + (size_t)k * h), x[c0], a0);
        a1 = fmaf(ld_stream(v + (size_t)(k + 1) * h), x[c1], a1);
    }
    if (k < w) a0 = fmaf(ld_stream(v + (size_t)k * h), x[ld_stream(c + (size_t)k * h)], a0);
    y[r] += a0 + a1;
}

}  // namespace

namespace launch {

hipError_t ell(hipStream_t s, int rows, int hal, int ellnum, int rows_per_lane,
               const int* col, const float* val, const float* x, float* y) {
    if (rows <= 0) return hipSuccess;
    if (rows_per_lane == 4) {
        if (hal % 4 != 0) return hipErrorInvalidValue;
        const int lanes = (rows + 3) / 4;
        hipLaunchKernelGGL(ell4_kernel, dim3((lanes + kBlock - 1) / kBlock),
                           dim3(kBlock), 0, s, rows, hal, ellnum, col, val, x, y);
    } else {
        hipLaunchKernelGGL(ell_kernel, dim3((rows + kBlock - 1) / kBlock),
                           dim3(kBlock), 0, s, rows, hal, ellnum, col, val, x, y);
    }
    return hipGetLastError();
}

hipError_t sell(hipStream_t s, int rows, int slice_height, const int* ptr,
                const int* col, const float* val, const float* x, float* y) {
    if (rows <= 0) return hipSuccess;
    if (slice_height <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(sell_kernel, dim3((rows + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, rows, slice_height, ptr, col, val, x, y);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace clspmv
