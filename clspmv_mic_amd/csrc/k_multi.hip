// Multi-right-hand-side SpMV for gfx950: Y[:, q] += A X[:, q] for q < k
// (k <= 16) in one pass over the matrix, for DIA and CSR.
//
// X is row-major (the k values of column c are X[c*k .. c*k+k)), so the
// gather of one column brings every right-hand side's value in one or a few
// 16-byte loads: a random column costs one cache line for all k products
// instead of one line per product.  CSR rows longer than kMultiLongRow get
// a work-group of their own.  Y is column-major (ldy), so each
// right-hand side's y stays a contiguous vector; lanes that own consecutive
// rows write consecutive addresses.  Matrix values may be FP32 or 16-bit
// (widened on load); sums are FP32 in a fixed order, no atomics.

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;

// Loads the k (<= KB) values of row c of X.
template <int KB>
__device__ __forceinline__ void load_xrow(const float* __restrict__ X, int c, int k, bool v4,
                                          float* xv) {
    const float* p = X + (size_t)c * k;
    if (v4) {  // k % 4 == 0 and X 16-byte aligned
#pragma unroll
        for (int q = 0; q < KB; q += 4) {
            if (q < k) {
                const float4 t = *reinterpret_cast<const float4*>(p + q);
                xv[q] = t.x;
                xv[q + 1] = t.y;
                xv[q + 2] = t.z;
                xv[q + 3] = t.w;
            } else {
                xv[q] = xv[q + 1] = xv[q + 2] = xv[q + 3] = 0.0f;
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < KB; ++q) xv[q] = q < k ? p[q] : 0.0f;
    }
}

template <int KB, class V>
__global__ __launch_bounds__(kBlock) void dia_multi_kernel(
    int rows, int cols, int len, int ndiag, const int* __restrict__ off,
    const typename V::T* __restrict__ data, int k, const float* __restrict__ X, bool v4,
    float* __restrict__ Y, int ldy) {
    const int r = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (r >= rows) return;
    float acc[KB];
#pragma unroll
    for (int q = 0; q < KB; ++q) acc[q] = 0.0f;
    for (int j = 0; j < ndiag; ++j) {
        const float d = V::ld(data + (size_t)j * len + r);
        const int c = r + off[j];
        if ((unsigned)c < (unsigned)cols) {
            float xv[KB];
            load_xrow<KB>(X, c, k, v4, xv);
#pragma unroll
            for (int q = 0; q < KB; ++q) acc[q] = fmaf(d, xv[q], acc[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < KB; ++q)
        if (q < k) Y[(size_t)q * ldy + r] += acc[q];
}

// G lanes per row stride over it; the per-lane partial sums of every
// right-hand side are then combined by an xor tree over the G lanes.
template <int G, int KB, class V>
__global__ __launch_bounds__(kBlock) void csr_multi_kernel(
    int rows, const int* __restrict__ rp, const int* __restrict__ col,
    const typename V::T* __restrict__ val, int k, const float* __restrict__ X, bool v4,
    float* __restrict__ Y, int ldy) {
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int row = (tile * kBlock + (int)threadIdx.x) / G;
    const int lane = threadIdx.x & (G - 1);
    if (row >= rows) return;  // a lane group exits together
    const int s = rp[row], e = rp[row + 1];
    if (e - s > launch::kMultiLongRow) return;  // csr_multi_long_kernel's row
    float acc[KB];
#pragma unroll
    for (int q = 0; q < KB; ++q) acc[q] = 0.0f;
    for (int j = s + lane; j < e; j += G) {
        const int c = ld_stream(col + j);
        const float v = V::ld(val + j);
        float xv[KB];
        load_xrow<KB>(X, c, k, v4, xv);
#pragma unroll
        for (int q = 0; q < KB; ++q) acc[q] = fmaf(v, xv[q], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < KB; ++q) acc[q] = group_sum<G>(acc[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < KB; ++q)
            if (q < k) Y[(size_t)q * ldy + row] += acc[q];
    }
}

// Rows longer than kMultiLongRow (power-law tails): one work-group per row,
// its four waves' partial sums combined in a fixed order through LDS.
template <int KB, class V>
__global__ __launch_bounds__(kBlock) void csr_multi_long_kernel(
    const int* __restrict__ long_rows, const int* __restrict__ rp, const int* __restrict__ col,
    const typename V::T* __restrict__ val, int k, const float* __restrict__ X, bool v4,
    float* __restrict__ Y, int ldy) {
    __shared__ float part[kBlock / kWave][KB];
    const int row = long_rows[blockIdx.x];
    const int t = threadIdx.x;
    const int s = rp[row], e = rp[row + 1];
    float acc[KB];
#pragma unroll
    for (int q = 0; q < KB; ++q) acc[q] = 0.0f;
    for (int j = s + t; j < e; j += kBlock) {
        const int c = ld_stream(col + j);
        const float v = V::ld(val + j);
        float xv[KB];
        load_xrow<KB>(X, c, k, v4, xv);
#pragma unroll
        for (int q = 0; q < KB; ++q) acc[q] = fmaf(v, xv[q], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < KB; ++q) acc[q] = group_sum<kWave>(acc[q]);
    if ((t & (kWave - 1)) == 0) {
#pragma unroll
        for (int q = 0; q < KB; ++q) part[t / kWave][q] = acc[q];
    }
    __syncthreads();
    if (t < KB && t < k) {
        float tot = 0.0f;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) tot += part[w][t];
        Y[(size_t)t * ldy + row] += tot;
    }
}

int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

template <int KB, class V>
hipError_t dia_multi_kb(hipStream_t s, int rows, int cols, int len, int ndiag, const int* off,
                        const void* data, int k, const float* X, bool v4, float* Y, int ldy) {
    hipLaunchKernelGGL((dia_multi_kernel<KB, V>), dim3(ceil_div(rows, kBlock)), dim3(kBlock), 0, s, rows,
                       cols, len, ndiag, off, static_cast<const typename V::T*>(data), k, X, v4, Y, ldy);
    return hipGetLastError();
}

template <class V>
hipError_t dia_multi_v(hipStream_t s, int rows, int cols, int len, int ndiag, const int* off,
                       const void* data, int k, const float* X, bool v4, float* Y, int ldy) {
    if (k <= 1) return dia_multi_kb<1, V>(s, rows, cols, len, ndiag, off, data, k, X, false, Y, ldy);
    if (k <= 4) return dia_multi_kb<4, V>(s, rows, cols, len, ndiag, off, data, k, X, v4, Y, ldy);
    if (k <= 8) return dia_multi_kb<8, V>(s, rows, cols, len, ndiag, off, data, k, X, v4, Y, ldy);
    return dia_multi_kb<16, V>(s, rows, cols, len, ndiag, off, data, k, X, v4, Y, ldy);
}

template <int G, int KB, class V>
hipError_t csr_multi_g(hipStream_t s, int rows, const int* rp, const int* col, const void* val, int k,
                       const float* X, bool v4, float* Y, int ldy) {
    hipLaunchKernelGGL((csr_multi_kernel<G, KB, V>), dim3(ceil_div((long long)rows * G, kBlock)),
                       dim3(kBlock), 0, s, rows, rp, col, static_cast<const typename V::T*>(val), k, X, v4,
                       Y, ldy);
    return hipGetLastError();
}

template <int KB, class V>
hipError_t csr_multi_kb(hipStream_t s, int rows, int lanes, const int* rp, const int* col, const void* val,
                        int k, const float* X, bool v4, float* Y, int ldy) {
    switch (lanes) {
        case 1: return csr_multi_g<1, KB, V>(s, rows, rp, col, val, k, X, v4, Y, ldy);
        case 2: return csr_multi_g<2, KB, V>(s, rows, rp, col, val, k, X, v4, Y, ldy);
        case 4: return csr_multi_g<4, KB, V>(s, rows, rp, col, val, k, X, v4, Y, ldy);
        case 8: return csr_multi_g<8, KB, V>(s, rows, rp, col, val, k, X, v4, Y, ldy);
        case 16: return csr_multi_g<16, KB, V>(s, rows, rp, col, val, k, X, v4, Y, ldy);
        case 32: return csr_multi_g<32, KB, V>(s, rows, rp, col, val, k, X, v4, Y, ldy);
        case 64: return csr_multi_g<64, KB, V>(s, rows, rp, col, val, k, X, v4, Y, ldy);
        default: return hipErrorInvalidValue;
    }
}

template <int KB, class V>
hipError_t csr_multi_long_kb(hipStream_t s, int nlong, const int* long_rows, const int* rp, const int* col,
                             const void* val, int k, const float* X, bool v4, float* Y, int ldy) {
    if (nlong <= 0) return hipSuccess;
    hipLaunchKernelGGL((csr_multi_long_kernel<KB, V>), dim3(nlong), dim3(kBlock), 0, s, long_rows, rp, col,
                       static_cast<const typename V::T*>(val), k, X, v4, Y, ldy);
    return hipGetLastError();
}

template <class V>
hipError_t csr_multi_v(hipStream_t s, int rows, int lanes, const int* rp, const int* col, const void* val,
                       int nlong, const int* long_rows, int k, const float* X, bool v4, float* Y, int ldy) {
    hipError_t e;
    if (k <= 1) {
        e = csr_multi_kb<1, V>(s, rows, lanes, rp, col, val, k, X, false, Y, ldy);
        if (e == hipSuccess) e = csr_multi_long_kb<1, V>(s, nlong, long_rows, rp, col, val, k, X, false, Y, ldy);
    } else if (k <= 4) {
        e = csr_multi_kb<4, V>(s, rows, lanes, rp, col, val, k, X, v4, Y, ldy);
        if (e == hipSuccess) e = csr_multi_long_kb<4, V>(s, nlong, long_rows, rp, col, val, k, X, v4, Y, ldy);
    } else if (k <= 8) {
        e = csr_multi_kb<8, V>(s, rows, lanes, rp, col, val, k, X, v4, Y, ldy);
        if (e == hipSuccess) e = csr_multi_long_kb<8, V>(s, nlong, long_rows, rp, col, val, k, X, v4, Y, ldy);
    } else {
        e = csr_multi_kb<16, V>(s, rows, lanes, rp, col, val, k, X, v4, Y, ldy);
        if (e == hipSuccess) e = csr_multi_long_kb<16, V>(s, nlong, long_rows, rp, col, val, k, X, v4, Y, ldy);
    }
    return e;
}

}  // namespace

namespace launch {

hipError_t dia_multi(hipStream_t s, int rows, int cols, int len, int ndiag, const int* offsets,
                     int dtype, const void* data, int k, const float* X, float* Y, int ldy) {
    if (rows <= 0 || ndiag <= 0 || k <= 0) return hipSuccess;
    if (k > kMultiMax) return hipErrorInvalidValue;
    const bool v4 = k % 4 == 0 && ((uintptr_t)X & 15) == 0;
    switch (dtype) {
        case 0: return dia_multi_v<ValF32>(s, rows, cols, len, ndiag, offsets, data, k, X, v4, Y, ldy);
        case 1: return dia_multi_v<ValBF16>(s, rows, cols, len, ndiag, offsets, data, k, X, v4, Y, ldy);
        case 2: return dia_multi_v<ValF16>(s, rows, cols, len, ndiag, offsets, data, k, X, v4, Y, ldy);
        default: return hipErrorInvalidValue;
    }
}

hipError_t csr_multi(hipStream_t s, int rows, int lanes, const int* rp, const int* col, int dtype,
                     const void* val, int nlong, const int* long_rows, int k, const float* X, float* Y,
                     int ldy) {
    if (rows <= 0 || k <= 0) return hipSuccess;
    if (k > kMultiMax) return hipErrorInvalidValue;
    const bool v4 = k % 4 == 0 && ((uintptr_t)X & 15) == 0;
    switch (dtype) {
        case 0: return csr_multi_v<ValF32>(s, rows, lanes, rp, col, val, nlong, long_rows, k, X, v4, Y, ldy);
        case 1: return csr_multi_v<ValBF16>(s, rows, lanes, rp, col, val, nlong, long_rows, k, X, v4, Y, ldy);
        case 2: return csr_multi_v<ValF16>(s, rows, lanes, rp, col, val, nlong, long_rows, k, X, v4, Y, ldy);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace launch
}  // namespace clspmv
