// Blocked SpMV for gfx950: BCSR (b4csr), BELL (b4ell) and SBELL layouts
// (replace kernels/spmv_bcsr.cl, kernels/spmv_bell.cl, kernels/spmv_sbell.cl).
//
// Block shapes bh x bw with bh in {1,2,4,8} and bw in {4,8}; every layout
// keeps each block's row-h, column-quad-w4 values as one 16-byte float4 and
// puts the float4s of neighbouring blocks (BCSR) or block rows (BELL/SBELL)
// next to each other, so a wave's loads of one (h, w4) plane are contiguous.
// Block columns are stored in float4 units: x is read as float4 too.
//
// BCSR: the reference spends a 128-lane work-group per block row
// (kernels/spmv_bcsr.cl:4-40); here a group of G lanes (G from the average
// number of blocks per block row) shares a block row, strides over its
// blocks and combines its bh partial sums with lane shuffles.
// BELL/SBELL: one lane per block row walks its block slots.

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;

template <int BH, int BW4, int G>
__global__ __launch_bounds__(kBlock) void bcsr_kernel(
    int brows, int rows, int cols, int aligned, const int* __restrict__ rp,
    const int* __restrict__ col4, const float* __restrict__ data,
    const float* __restrict__ x, bool x_aligned, float* __restrict__ y) {
    const int g = (xcd_tile(blockIdx.x, gridDim.x) * kBlock + (int)threadIdx.x) / G;
    const int lane = threadIdx.x & (G - 1);
    if (g >= brows) return;  // lane groups exit together
    const int s = rp[g], e = rp[g + 1];
    float acc[BH];
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = 0.0f;
    const size_t plane = (size_t)aligned * BH;  // floats per w4 plane
    for (int j = s + lane; j < e; j += G) {
        const int c4 = ld_stream(col4 + j);
#pragma unroll
        for (int w4 = 0; w4 < BW4; ++w4) {
            const float4 xv = ld_x4(x, (c4 + w4) * 4, cols, x_aligned);
#pragma unroll
            for (int h = 0; h < BH; ++h) {
                const float4 d = ld_stream(reinterpret_cast<const float4*>(
                    data + w4 * plane + (size_t)h * aligned + (size_t)j * 4));
                acc[h] = dot4(d, xv, acc[h]);
            }
        }
    }
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = group_sum<G>(acc[h]);
    // Lane h % G of the group writes output row h.
#pragma unroll
    for (int h = 0; h < BH; ++h) {
        const int r = g * BH + h;
        if (h % G == lane && r < rows) y[r] += acc[h];
    }
}

// BCSR in the slot-major encoding: one lane per block row walks its slots,
// each slot's block columns and value planes contiguous across the wave
// (BELL's pattern).  Padding slots (column -1) hold zero values; their loads
// go to block column 0 and their products are not added, so the sum is the
// row's blocks' in block order, one float4 dot per (w4, h).
template <int BH, int BW4>
__global__ __launch_bounds__(kBlock) void bcsr_slots_kernel(
    int brows, int rows, int cols, int hal, int nslots, const int* __restrict__ col4,
    const float* __restrict__ data, const float* __restrict__ x, bool x_aligned, float* __restrict__ y) {
    const int r = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (r >= brows) return;
    float acc[BH];
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = 0.0f;
    const size_t plane = (size_t)hal * 4;  // floats per (slot, w4, h) plane
    for (int sl = 0; sl < nslots; ++sl) {
        const int c4 = ld_stream(col4 + (size_t)sl * hal + r);
        const bool live = c4 >= 0;
        const float* base = data + (size_t)sl * BW4 * BH * plane + (size_t)r * 4;
#pragma unroll
        for (int w4 = 0; w4 < BW4; ++w4) {
            const float4 xv = ld_x4(x, ((live ? c4 : 0) + w4) * 4, cols, x_aligned);
#pragma unroll
            for (int h = 0; h < BH; ++h) {
                const float4 d = ld_stream(reinterpret_cast<const float4*>(base + (w4 * BH + h) * plane));
                const float a = dot4(d, xv, acc[h]);
                acc[h] = live ? a : acc[h];
            }
        }
    }
#pragma unroll
    for (int h = 0; h < BH; ++h) {
        const int row = r * BH + h;
        if (row < rows) y[row] += acc[h];
    }
}

// F4ACC = false: one scalar accumulator per block row row (method 0);
// true: float4 accumulators reduced at the end (method 1, the reference's
// "_mad" variant, kernels/spmv_bell.cl:24-42).
template <int BH, int BW4, bool F4ACC>
__global__ __launch_bounds__(kBlock) void bell_kernel(
    int brows, int rows, int cols, int hal, int f4, int ellnum,
    const int* __restrict__ col4, const float* __restrict__ data,
    const float* __restrict__ x, bool x_aligned, float* __restrict__ y) {
    const int r = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (r >= brows) return;
    float acc[BH];
    float4 acc4[BH];
#pragma unroll
    for (int h = 0; h < BH; ++h) {
        acc[h] = 0.0f;
        acc4[h] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const size_t slot_stride = (size_t)f4 * BW4 * BH;
    const size_t w4_stride = (size_t)f4 * BH;
    for (int sl = 0; sl < ellnum; ++sl) {
        const int c4 = ld_stream(col4 + r + (size_t)sl * hal);
        const float* base = data + sl * slot_stride + (size_t)r * 4;
#pragma unroll
        for (int w4 = 0; w4 < BW4; ++w4) {
            const float4 xv = ld_x4(x, (c4 + w4) * 4, cols, x_aligned);
#pragma unroll
            for (int h = 0; h < BH; ++h) {
                const float4 d = ld_stream(reinterpret_cast<const float4*>(
                    base + (size_t)h * f4 + w4 * w4_stride));
                if (F4ACC) {
                    acc4[h].x = fmaf(d.x, xv.x, acc4[h].x);
                    acc4[h].y = fmaf(d.y, xv.y, acc4[h].y);
                    acc4[h].z = fmaf(d.z, xv.z, acc4[h].z);
                    acc4[h].w = fmaf(d.w, xv.w, acc4[h].w);
                } else {
                    acc[h] = dot4(d, xv, acc[h]);
                }
            }
        }
    }
#pragma unroll
    for (int h = 0; h < BH; ++h) {
        const int row = r * BH + h;
        const float v = F4ACC ? (acc4[h].x + acc4[h].y) + (acc4[h].z + acc4[h].w) : acc[h];
        if (row < rows) y[row] += v;
    }
}

template <int BH, int BW4>
__global__ __launch_bounds__(kBlock) void sbell_kernel(
    int brows, int rows, int cols, int sh, const int* __restrict__ sptr,
    const int* __restrict__ col4, const float* __restrict__ data,
    const float* __restrict__ x, bool x_aligned, float* __restrict__ y) {
    const int r = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (r >= brows) return;
    const int sl = r / sh, i = r - sl * sh;
    const int p0 = sptr[sl];
    const int w = (sptr[sl + 1] - p0) / sh;
    const float* base = data + (size_t)p0 * BH * BW4 * 4 + (size_t)i * 4;
    const size_t slot_stride = (size_t)sh * BW4 * BH * 4;
    const size_t h_stride = (size_t)sh * 4;
    const size_t w4_stride = (size_t)sh * BH * 4;
    float acc[BH];
#pragma unroll
    for (int h = 0; h < BH; ++h) acc[h] = 0.0f;
    for (int k = 0; k < w; ++k) {
        const int c4 = ld_stream(col4 + p0 + i + (size_t)k * sh);
#pragma unroll
        for (int w4 = 0; w4 < BW4; ++w4) {
            const float4 xv = ld_x4(x, (c4 + w4) * 4, cols, x_aligned);
#pragma unroll
            for (int h = 0; h < BH; ++h) {
                const float4 d = ld_stream(reinterpret_cast<const float4*>(
                    base + k * slot_stride + h * h_stride + w4 * w4_stride));
                acc[h] = dot4(d, xv, acc[h]);
            }
        }
    }
#pragma unroll
    for (int h = 0; h < BH; ++h) {
        const int row = r * BH + h;
        if (row < rows) y[row] += acc[h];
    }
}

int grid_for(long long threads) { return (int)((threads + kBlock - 1) / kBlock); }

template <int BH, int BW4>
hipError_t bcsr_shape(hipStream_t s, int brows, int rows, int cols, int lanes,
                      int aligned, const int* rp, const int* col4,
                      const float* data, const float* x, bool xa, float* y) {
    const int grid = grid_for((long long)brows * lanes);
#define BCSRG(G)                                                                    \
    case G:                                                                         \
        hipLaunchKernelGGL((bcsr_kernel<BH, BW4, G>), dim3(grid), dim3(kBlock), 0, s, \
                           brows, rows, cols, aligned, rp, col4, data, x, xa, y);   \
        break;
    switch (lanes) {
        BCSRG(1) BCSRG(2) BCSRG(4) BCSRG(8) BCSRG(16) BCSRG(32) BCSRG(64)
        default: return hipErrorInvalidValue;
    }
#undef BCSRG
    return hipGetLastError();
}

}  // namespace

namespace launch {

#define SHAPE_DISPATCH(FN, ...)                                   \
    switch (bh * 10 + bw) {                                       \
        case 14: return FN<1, 1>(__VA_ARGS__);                    \
        case 24: return FN<2, 1>(__VA_ARGS__);                    \
        case 44: return FN<4, 1>(__VA_ARGS__);                    \
        case 84: return FN<8, 1>(__VA_ARGS__);                    \
        case 18: return FN<1, 2>(__VA_ARGS__);                    \
        case 28: return FN<2, 2>(__VA_ARGS__);                    \
        case 48: return FN<4, 2>(__VA_ARGS__);                    \
        case 88: return FN<8, 2>(__VA_ARGS__);                    \
        default: return hipErrorInvalidValue;                     \
    }

hipError_t bcsr(hipStream_t s, int brows, int rows, int cols, int bh, int bw,
                int lanes, int aligned, const int* rp, const int* col4,
                const float* data, const float* x, bool x_aligned, float* y) {
    if (brows <= 0) return hipSuccess;
    SHAPE_DISPATCH(bcsr_shape, s, brows, rows, cols, lanes, aligned, rp, col4,
                   data, x, x_aligned, y)
}

template <int BH, int BW4>
static hipError_t bcsr_slots_shape(hipStream_t s, int brows, int rows, int cols, int hal, int nslots,
                                   const int* col4, const float* data, const float* x, bool xa, float* y) {
    hipLaunchKernelGGL((bcsr_slots_kernel<BH, BW4>), dim3(grid_for(brows)), dim3(kBlock), 0, s, brows, rows, cols,
                       hal, nslots, col4, data, x, xa, y);
    return hipGetLastError();
}

hipError_t bcsr_slots(hipStream_t s, int brows, int rows, int cols, int bh, int bw, int hal, int nslots,
                      const int* col4, const float* data, const float* x, bool x_aligned, float* y) {
    if (brows <= 0 || nslots <= 0) return hipSuccess;
    if (hal < brows || nslots > kBcsrSlotsMax) return hipErrorInvalidValue;
    SHAPE_DISPATCH(bcsr_slots_shape, s, brows, rows, cols, hal, nslots, col4, data, x, x_aligned, y)
}

template <int BH, int BW4>
static hipError_t bell_shape(hipStream_t s, int brows, int rows, int cols,
                             int method, int hal, int f4, int ellnum,
                             const int* col4, const float* data, const float* x,
                             bool xa, float* y) {
    const int grid = grid_for(brows);
    if (method == 1)
        hipLaunchKernelGGL((bell_kernel<BH, BW4, true>), dim3(grid), dim3(kBlock), 0, s,
                           brows, rows, cols, hal, f4, ellnum, col4, data, x, xa, y);
    else
        hipLaunchKernelGGL((bell_kernel<BH, BW4, false>), dim3(grid), dim3(kBlock), 0, s,
                           brows, rows, cols, hal, f4, ellnum, col4, data, x, xa, y);
    return hipGetLastError();
}

hipError_t bell(hipStream_t s, int brows, int rows, int cols, int bh, int bw,
                int method, int hal, int f4, int ellnum, const int* col4,
                const float* data, const float* x, bool x_aligned, float* y) {
    if (brows <= 0 || ellnum <= 0) return hipSuccess;
    SHAPE_DISPATCH(bell_shape, s, brows, rows, cols, method, hal, f4, ellnum,
                   col4, data, x, x_aligned, y)
}

template <int BH, int BW4>
static hipError_t sbell_shape(hipStream_t s, int brows, int rows, int cols,
                              int sh, const int* sptr, const int* col4,
                              const float* data, const float* x, bool xa, float* y) {
    hipLaunchKernelGGL((sbell_kernel<BH, BW4>), dim3(grid_for(brows)), dim3(kBlock), 0, s,
                       brows, rows, cols, sh, sptr, col4, data, x, xa, y);
    return hipGetLastError();
}

hipError_t sbell(hipStream_t s, int brows, int rows, int cols, int bh, int bw,
                 int slice_height, const int* sptr, const int* col4,
                 const float* data, const float* x, bool x_aligned, float* y) {
    if (brows <= 0) return hipSuccess;
    if (slice_height <= 0) return hipErrorInvalidValue;
    SHAPE_DISPATCH(sbell_shape, s, brows, rows, cols, slice_height, sptr, col4,
                   data, x, x_aligned, y)
}

#undef SHAPE_DISPATCH

}  // namespace launch
}  // namespace clspmv
