// DIA and BDIA SpMV for gfx950 (replace kernels/spmv_dia.cl and
// kernels/spmv_bdia.cl).
//
// Storage is diagonal-major: diagonal j of row r sits at data[j*len + r], so
// each diagonal is one coalesced stream.  Offsets are wave-uniform and come
// in through scalar loads.  x is read either directly (every row's x
// elements are contiguous across the wave and reused from L1/L2) or from an
// LDS window staged once per work-group:
//   DIA : one window covering [min offset, max offset + rows of the group)
//   BDIA: one window per band (consecutive offsets), double buffered so that
//         staging band b+1 needs a single barrier.
// Out-of-range columns read zero, which replaces the reference's padded
// `vec_extended` copy of x (src/spmv_cocktail.cpp:253-301).  Unlike the
// reference's gpu_bdia there is no band-count or band-width limit
// (MAX_BAND_NUM 31, MAX_BAND_WIDTH 128); bands wider than the LDS window fall
// back to direct x loads.

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;

__device__ __forceinline__ float x_at(const float* __restrict__ x, int c, int cols) {
    return (unsigned)c < (unsigned)cols ? x[c] : 0.0f;
}

// ---- DIA ---------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void dia1_kernel(
    int rows, int cols, int len, int ndiag, const int* __restrict__ off,
    const float* __restrict__ data, const float* __restrict__ x,
    float* __restrict__ y) {
    const int r = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (r >= rows) return;
    float a0 = 0.0f, a1 = 0.0f;
    int j = 0;
    for (; j + 1 < ndiag; j += 2) {
        a0 = fmaf(ld_stream(data + (size_t)j * len + r), x_at(x, r + off[j], cols), a0);
        a1 = fmaf(ld_stream(data + (size_t)(j + 1) * len + r), x_at(x, r + off[j + 1], cols), a1);
    }
    if (j < ndiag) a0 = fmaf(ld_stream(data + (size_t)j * len + r), x_at(x, r + off[j], cols), a0);
    y[r] += a0 + a1;
}

// Four consecutive rows per lane, 128-bit data loads (len % 4 == 0).
__global__ __launch_bounds__(kBlock) void dia4_kernel(
    int rows, int cols, int len, int ndiag, const int* __restrict__ off,
    const float* __restrict__ data, const float* __restrict__ x,
    float* __restrict__ y) {
    const int r = (xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x) * 4;
    if (r >= rows) return;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int j = 0; j < ndiag; ++j) {
        const float4 d = ld_stream(reinterpret_cast<const float4*>(data + (size_t)j * len + r));
        const int c = r + off[j];
        a.x = fmaf(d.x, x_at(x, c + 0, cols), a.x);
        a.y = fmaf(d.y, x_at(x, c + 1, cols), a.y);
        a.z = fmaf(d.z, x_at(x, c + 2, cols), a.z);
        a.w = fmaf(d.w, x_at(x, c + 3, cols), a.w);
    }
    y[r] += a.x;
    if (r + 1 < rows) y[r + 1] += a.y;
    if (r + 2 < rows) y[r + 2] += a.z;
    if (r + 3 < rows) y[r + 3] += a.w;
}

// RPT rows per lane (strided by the work-group size) with the work-group's x
// window in LDS; the host guarantees RPT*kBlock + span <= kLdsWindow.
template <int RPT>
__global__ __launch_bounds__(kBlock) void dia_lds_kernel(
    int rows, int cols, int len, int ndiag, const int* __restrict__ off,
    int min_off, int span, const float* __restrict__ data,
    const float* __restrict__ x, float* __restrict__ y) {
    __shared__ float win[launch::kLdsWindow];
    const int t = threadIdx.x;
    const int r0 = xcd_tile(blockIdx.x, gridDim.x) * kBlock * RPT;
    const int ws = r0 + min_off;
    const int wn = kBlock * RPT + span;
    for (int i = t; i < wn; i += kBlock) win[i] = x_at(x, ws + i, cols);
    __syncthreads();
    float a[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) a[q] = 0.0f;
    for (int j = 0; j < ndiag; ++j) {
        const int o = off[j] - min_off;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int rl = q * kBlock + t;
            if (r0 + rl < rows)
                a[q] = fmaf(ld_stream(data + (size_t)j * len + r0 + rl), win[rl + o], a[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = r0 + q * kBlock + t;
        if (r < rows) y[r] += a[q];
    }
}

// ---- BDIA --------------------------------------------------------------------

template <int RPT>
__global__ __launch_bounds__(kBlock) void bdia_direct_kernel(
    int rows, int cols, int len, int nband, const int* __restrict__ boff,
    const int* __restrict__ bptr, const float* __restrict__ data,
    const float* __restrict__ x, float* __restrict__ y) {
    // RPT == 1: one row per lane; RPT == 4: four consecutive rows per lane
    // with 128-bit data loads (len % 4 == 0).
    const int r = (xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x) * RPT;
    if (r >= rows) return;
    float a[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) a[q] = 0.0f;
    for (int b = 0; b < nband; ++b) {
        const int d0 = bptr[b], d1 = bptr[b + 1];
        const int o = boff[b] - d0;  // offset of diagonal d is o + d
        for (int d = d0; d < d1; ++d) {
            const float* dp = data + (size_t)d * len + r;
            const int c = r + o + d;
            if (RPT == 4) {
                const float4 v = ld_stream(reinterpret_cast<const float4*>(dp));
                a[0] = fmaf(v.x, x_at(x, c + 0, cols), a[0]);
                a[1 % RPT] = fmaf(v.y, x_at(x, c + 1, cols), a[1 % RPT]);
                a[2 % RPT] = fmaf(v.z, x_at(x, c + 2, cols), a[2 % RPT]);
                a[3 % RPT] = fmaf(v.w, x_at(x, c + 3, cols), a[3 % RPT]);
            } else {
                a[0] = fmaf(ld_stream(dp), x_at(x, c, cols), a[0]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q)
        if (r + q < rows) y[r + q] += a[q];
}

// RPT rows per lane strided by the work-group size; per band an LDS window
// [r0 + first offset, r0 + first offset + RPT*kBlock + width - 1), double
// buffered.  The host guarantees RPT*kBlock + max band width <= kLdsWindow/2.
template <int RPT>
__global__ __launch_bounds__(kBlock) void bdia_lds_kernel(
    int rows, int cols, int len, int nband, const int* __restrict__ boff,
    const int* __restrict__ bptr, const float* __restrict__ data,
    const float* __restrict__ x, float* __restrict__ y) {
    constexpr int kHalf = launch::kLdsWindow / 2;
    __shared__ float win[2][kHalf];
    const int t = threadIdx.x;
    const int r0 = xcd_tile(blockIdx.x, gridDim.x) * kBlock * RPT;
    float a[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) a[q] = 0.0f;
    for (int b = 0; b < nband; ++b) {
        const int d0 = bptr[b], d1 = bptr[b + 1];
        const int ws = r0 + boff[b];
        const int wn = kBlock * RPT + (d1 - d0) - 1;
        float* w = win[b & 1];
        for (int i = t; i < wn; i += kBlock) w[i] = x_at(x, ws + i, cols);
        __syncthreads();
        for (int d = d0; d < d1; ++d) {
            const int k = d - d0;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int rl = q * kBlock + t;
                if (r0 + rl < rows)
                    a[q] = fmaf(ld_stream(data + (size_t)d * len + r0 + rl), w[rl + k], a[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = r0 + q * kBlock + t;
        if (r < rows) y[r] += a[q];
    }
}

int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace

namespace launch {

hipError_t dia(hipStream_t s, int rows, int cols, int len, int ndiag,
               const int* offsets, int min_off, int max_off, int method,
               const float* data, const float* x, float* y) {
    if (rows <= 0 || ndiag <= 0) return hipSuccess;
    const int span = max_off - min_off;
    switch (method) {
        case 0:
            hipLaunchKernelGGL(dia1_kernel, dim3(ceil_div(rows, kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, ndiag, offsets, data, x, y);
            break;
        case 1:
            if (len % 4) return hipErrorInvalidValue;
            hipLaunchKernelGGL(dia4_kernel, dim3(ceil_div(rows, 4 * kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, ndiag, offsets, data, x, y);
            break;
        case 2:
            if (kBlock + span > kLdsWindow) return hipErrorInvalidValue;
            hipLaunchKernelGGL(dia_lds_kernel<1>, dim3(ceil_div(rows, kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, ndiag, offsets, min_off, span, data, x, y);
            break;
        case 3:
            if (4 * kBlock + span > kLdsWindow) return hipErrorInvalidValue;
            hipLaunchKernelGGL(dia_lds_kernel<4>, dim3(ceil_div(rows, 4 * kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, ndiag, offsets, min_off, span, data, x, y);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t bdia(hipStream_t s, int rows, int cols, int len, int nband,
                const int* band_off, const int* bptr, int max_band_width,
                int method, const float* data, const float* x, float* y) {
    if (rows <= 0 || nband <= 0) return hipSuccess;
    const int half = kLdsWindow / 2;
    switch (method) {
        case 0:
            hipLaunchKernelGGL(bdia_direct_kernel<1>, dim3(ceil_div(rows, kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, nband, band_off, bptr, data, x, y);
            break;
        case 1:
            if (kBlock + max_band_width > half) return hipErrorInvalidValue;
            hipLaunchKernelGGL(bdia_lds_kernel<1>, dim3(ceil_div(rows, kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, nband, band_off, bptr, data, x, y);
            break;
        case 2:
            if (2 * kBlock + max_band_width > half) return hipErrorInvalidValue;
            hipLaunchKernelGGL(bdia_lds_kernel<2>, dim3(ceil_div(rows, 2 * kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, nband, band_off, bptr, data, x, y);
            break;
        case 3:
            if (len % 4) return hipErrorInvalidValue;
            hipLaunchKernelGGL(bdia_direct_kernel<4>, dim3(ceil_div(rows, 4 * kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, nband, band_off, bptr, data, x, y);
            break;
        case 4:
            if (4 * kBlock + max_band_width > half) return hipErrorInvalidValue;
            hipLaunchKernelGGL(bdia_lds_kernel<4>, dim3(ceil_div(rows, 4 * kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, nband, band_off, bptr, data, x, y);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace launch
}  // namespace clspmv
