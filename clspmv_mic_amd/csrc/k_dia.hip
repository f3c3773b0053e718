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
// BDIA methods without LDS run the DIA kernels over the band's diagonals
// listed one by one (the runtime expands the band offsets at upload).
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

// One row per lane; U diagonals per iteration with independent
// accumulators so that U data streams and U x reads are in flight per lane.
template <int U, class V>
__global__ __launch_bounds__(kBlock) void dia1_kernel(
    int rows, int cols, int len, int ndiag, const int* __restrict__ off,
    const typename V::T* __restrict__ data, const float* __restrict__ x,
    float* __restrict__ y) {
    const int r = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (r >= rows) return;
    float a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = 0.0f;
    int j = 0;
    for (; j + U <= ndiag; j += U) {
        float d[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) d[u] = V::ld(data + (size_t)(j + u) * len + r);
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x_at(x, r + off[j + u], cols);
#pragma unroll
        for (int u = 0; u < U; ++u) a[u] = fmaf(d[u], xv[u], a[u]);
    }
    for (; j < ndiag; ++j) a[0] = fmaf(V::ld(data + (size_t)j * len + r), x_at(x, r + off[j], cols), a[0]);
    float sum = 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u) sum += a[u];
    y[r] += sum;
}

// One row per lane with every diagonal in flight at once (ndiag <= N): the
// lane issues all N data loads and x loads before its first product, so a
// wave keeps N coalesced 256-byte streams outstanding instead of U, and the
// products are then summed in diagonal order.  Slots j >= ndiag re-load the
// last diagonal (cache hits) and are not added.
template <int N, class V>
__global__ __launch_bounds__(kBlock) void dia_row_kernel(
    int rows, int cols, int len, int ndiag, const int* __restrict__ off,
    const typename V::T* __restrict__ data, const float* __restrict__ x,
    float* __restrict__ y) {
    const int r = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (r >= rows) return;
    float d[N], xv[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int jj = j < ndiag ? j : ndiag - 1;  // wave-uniform
        d[j] = V::ld(data + (size_t)jj * len + r);
        xv[j] = x_at(x, r + off[jj], cols);
    }
    float a = 0.0f;
#pragma unroll
    for (int j = 0; j < N; ++j) a = j < ndiag ? fmaf(d[j], xv[j], a) : a;
    y[r] += a;
}

// Two consecutive rows per lane (64-bit data loads, len % 2 == 0), U
// diagonals per iteration.
template <int U, class V>
__global__ __launch_bounds__(kBlock) void dia2_kernel(
    int rows, int cols, int len, int ndiag, const int* __restrict__ off,
    const typename V::T* __restrict__ data, const float* __restrict__ x,
    float* __restrict__ y) {
    const int r = (xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x) * 2;
    if (r >= rows) return;
    float a0[U], a1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a0[u] = a1[u] = 0.0f;
    int j = 0;
    for (; j + U <= ndiag; j += U) {
        float2 d[U];
        float x0[U], x1[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            d[u] = V::ld2(data + (size_t)(j + u) * len + r);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = r + off[j + u];
            x0[u] = x_at(x, c, cols);
            x1[u] = x_at(x, c + 1, cols);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a0[u] = fmaf(d[u].x, x0[u], a0[u]);
            a1[u] = fmaf(d[u].y, x1[u], a1[u]);
        }
    }
    for (; j < ndiag; ++j) {
        const float2 d = V::ld2(data + (size_t)j * len + r);
        const int c = r + off[j];
        a0[0] = fmaf(d.x, x_at(x, c, cols), a0[0]);
        a1[0] = fmaf(d.y, x_at(x, c + 1, cols), a1[0]);
    }
    float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        s0 += a0[u];
        s1 += a1[u];
    }
    y[r] += s0;
    if (r + 1 < rows) y[r + 1] += s1;
}

// RPT rows per lane (strided by the work-group size) with the work-group's x
// window in LDS; the host guarantees RPT*kBlock + span <= kLdsWindow.
template <int RPT, class V>
__global__ __launch_bounds__(kBlock) void dia_lds_kernel(
    int rows, int cols, int len, int ndiag, const int* __restrict__ off,
    int min_off, int span, const typename V::T* __restrict__ data,
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
                a[q] = fmaf(V::ld(data + (size_t)j * len + r0 + rl), win[rl + o], a[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = r0 + q * kBlock + t;
        if (r < rows) y[r] += a[q];
    }
}

// ---- BDIA --------------------------------------------------------------------

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

template <class V>
hipError_t dia_typed(hipStream_t s, int rows, int cols, int len, int ndiag,
                     const int* offsets, int min_off, int max_off, int method,
                     const typename V::T* data, const float* x, float* y) {
    const int span = max_off - min_off;
    switch (method) {
        case 0: {
            // Up to 32 diagonals: all in flight (the smallest multiple of 4
            // slots that holds them); more: 8 at a time.
            const dim3 g(ceil_div(rows, kBlock)), b(kBlock);
            switch ((ndiag + 3) / 4) {
#define CLSPMV_DIA_ROW(Q)                                                                              \
    case Q:                                                                                           \
        hipLaunchKernelGGL((dia_row_kernel<4 * Q, V>), g, b, 0, s, rows, cols, len, ndiag, offsets, data, x, y); \
        break;
                CLSPMV_DIA_ROW(1)
                CLSPMV_DIA_ROW(2)
                CLSPMV_DIA_ROW(3)
                CLSPMV_DIA_ROW(4)
                CLSPMV_DIA_ROW(5)
                CLSPMV_DIA_ROW(6)
                CLSPMV_DIA_ROW(7)
                CLSPMV_DIA_ROW(8)
#undef CLSPMV_DIA_ROW
                default:
                    hipLaunchKernelGGL((dia1_kernel<8, V>), g, b, 0, s, rows, cols, len, ndiag, offsets, data, x, y);
            }
            break;
        }
        case 1:
            if (len % 2) return hipErrorInvalidValue;
            hipLaunchKernelGGL((dia2_kernel<4, V>), dim3(ceil_div(rows, 2 * kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, ndiag, offsets, data, x, y);
            break;
        case 2:
            if (kBlock + span > kLdsWindow) return hipErrorInvalidValue;
            hipLaunchKernelGGL((dia_lds_kernel<1, V>), dim3(ceil_div(rows, kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, ndiag, offsets, min_off, span, data, x, y);
            break;
        case 3:
            if (4 * kBlock + span > kLdsWindow) return hipErrorInvalidValue;
            hipLaunchKernelGGL((dia_lds_kernel<4, V>), dim3(ceil_div(rows, 4 * kBlock)), dim3(kBlock), 0, s,
                               rows, cols, len, ndiag, offsets, min_off, span, data, x, y);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t dia(hipStream_t s, int rows, int cols, int len, int ndiag,
               const int* offsets, int min_off, int max_off, int method, int dtype,
               const void* data, const float* x, float* y) {
    if (rows <= 0 || ndiag <= 0) return hipSuccess;
    switch (dtype) {
        case 0:
            return dia_typed<ValF32>(s, rows, cols, len, ndiag, offsets, min_off, max_off, method,
                                     static_cast<const float*>(data), x, y);
        case 1:
            return dia_typed<ValBF16>(s, rows, cols, len, ndiag, offsets, min_off, max_off, method,
                                      static_cast<const unsigned short*>(data), x, y);
        case 2:
            return dia_typed<ValF16>(s, rows, cols, len, ndiag, offsets, min_off, max_off, method,
                                     static_cast<const unsigned short*>(data), x, y);
        default:
            return hipErrorInvalidValue;
    }
}

hipError_t bdia(hipStream_t s, int rows, int cols, int len, int nband,
                const int* band_off, const int* bptr, int max_band_width,
                int method, const float* data, const float* x, float* y) {
    if (rows <= 0 || nband <= 0) return hipSuccess;
    const int half = kLdsWindow / 2;
    switch (method) {
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
