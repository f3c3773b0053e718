// CSR SpMV kernels for gfx950 (replace kernels/spmv_csr_scalar.cl and
// kernels/spmv_csr_vector.cl).
//
// csr_vector<G>: G lanes of a wave share one row (G = 1 is the scalar
//   kernel).  Lanes stride over the row, then a lane-group xor-shuffle sum
//   replaces the reference's barrier-free "volatile" LDS tree
//   (kernels/spmv_csr_vector.cl:32-38).
// csr_adaptive: one work-group per host-built row block.  Short rows: the
//   block's products val*x[col] are staged in LDS with coalesced loads, then
//   each row is summed by a lane group sized to the block's row count.  A row
//   longer than the LDS budget gets the whole work-group.  Every row belongs
//   to exactly one block, so y is written without atomics and results do not
//   depend on what else runs on the GPU.

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;

template <int G>
__global__ __launch_bounds__(kBlock) void csr_vector_kernel(
    int rows, const int* __restrict__ rp, const int* __restrict__ col,
    const float* __restrict__ val, const float* __restrict__ x,
    float* __restrict__ y) {
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int row = (tile * kBlock + (int)threadIdx.x) / G;
    const int lane = threadIdx.x & (G - 1);
    if (row >= rows) return;  // a lane group always exits together
    const int s = rp[row], e = rp[row + 1];
    float a0 = 0.0f, a1 = 0.0f;
    int k = s + lane;
    // Two independent accumulators keep two gathers in flight per lane.
    for (; k + G < e; k += 2 * G) {
        const int c0 = ld_stream(col + k), c1 = ld_stream(col + k + G);
        const float v0 = ld_stream(val + k), v1 = ld_stream(val + k + G);
        a0 = fmaf(v0, x[c0], a0);
        a1 = fmaf(v1, x[c1], a1);
    }
    if (k < e) a0 = fmaf(ld_stream(val + k), x[ld_stream(col + k)], a0);
    float sum = group_sum<G>(a0 + a1);
    if (lane == 0) y[row] += sum;
}

__global__ __launch_bounds__(kBlock) void csr_adaptive_kernel(
    const int* __restrict__ blocks, const int* __restrict__ rp,
    const int* __restrict__ col, const float* __restrict__ val,
    const float* __restrict__ x, float* __restrict__ y) {
    __shared__ float prod[launch::kAdaptiveNnz];
    __shared__ float wave_part[kBlock / kWave];
    const int b = blockIdx.x;
    const int r0 = blocks[b], r1 = blocks[b + 1];
    const int s = rp[r0], e = rp[r1];
    const int nrows = r1 - r0;
    const int t = threadIdx.x;

    constexpr int U = launch::kAdaptiveNnz / kBlock;  // nonzeros in flight per thread
    if (nrows == 1 && e - s > launch::kAdaptiveNnz) {
        // Long row: the whole work-group strides over it, U nonzeros' loads
        // in flight per thread (each thread's sum still in index order).
        float acc = 0.0f;
        int k = s + t;
        for (; k + (U - 1) * kBlock < e; k += U * kBlock) {
            int c[U];
            float v[U], xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c[u] = ld_stream(col + k + u * kBlock);
                v[u] = ld_stream(val + k + u * kBlock);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) acc = fmaf(v[u], xv[u], acc);
        }
        for (; k < e; k += kBlock) acc = fmaf(ld_stream(val + k), x[ld_stream(col + k)], acc);
        acc = group_sum<kWave>(acc);
        if ((t & (kWave - 1)) == 0) wave_part[t / kWave] = acc;
        __syncthreads();
        if (t == 0) {
            float tot = 0.0f;
#pragma unroll
            for (int w = 0; w < kBlock / kWave; ++w) tot += wave_part[w];
            y[r0] += tot;
        }
        return;
    }

    // Stage products with coalesced loads, every load of the block in
    // flight at once (e - s <= kAdaptiveNnz = U * kBlock).  A block of empty
    // rows loads nothing (its s may be the matrix's last index + 1).
    if (e > s) {
        int c[U];
        float v[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = t + u * kBlock;
            const int kk = k < e - s ? s + k : s;  // past the block: a valid index, not stored
            c[u] = ld_stream(col + kk);
            v[u] = ld_stream(val + kk);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (t + u * kBlock < e - s) prod[t + u * kBlock] = v[u] * xv[u];
    }
    __syncthreads();

    // Lane-group size: the largest power of two with nrows groups fitting
    // in the work-group, at most one wave.  A block holds at most
    // kAdaptiveRows (= kBlock) rows, so kBlock / g >= nrows and one pass
    // covers every row; all lanes run the shuffles.
    int g = kBlock / (nrows > 0 ? nrows : 1);
    g = g >= 64 ? 64 : g >= 32 ? 32 : g >= 16 ? 16 : g >= 8 ? 8 : g >= 4 ? 4
      : g >= 2 ? 2 : 1;
    const int lane = t & (g - 1);
    const int r = t / g;
    const bool live = r < nrows;
    float acc = 0.0f;
    if (live) {
        const int a = rp[r0 + r] - s, z = rp[r0 + r + 1] - s;
        for (int k = a + lane; k < z; k += g) acc += prod[k];
    }
    acc = group_sum_rt(acc, g);
    if (live && lane == 0) y[r0 + r] += acc;
}

}  // namespace

namespace launch {

hipError_t csr_vector(hipStream_t s, int rows, int lanes, const int* rp,
                      const int* col, const float* val, const float* x, float* y) {
    if (rows <= 0) return hipSuccess;
    const long long threads = (long long)rows * lanes;
    const int grid = (int)((threads + kBlock - 1) / kBlock);
#define CSRV(G)                                                                 \
    case G:                                                                     \
        hipLaunchKernelGGL(csr_vector_kernel<G>, dim3(grid), dim3(kBlock), 0, s, \
                           rows, rp, col, val, x, y);                           \
        break;
    switch (lanes) {
        CSRV(1) CSRV(2) CSRV(4) CSRV(8) CSRV(16) CSRV(32) CSRV(64)
        default: return hipErrorInvalidValue;
    }
#undef CSRV
    return hipGetLastError();
}

hipError_t csr_adaptive(hipStream_t s, int nblocks, const int* blocks,
                        const int* rp, const int* col, const float* val,
                        const float* x, float* y) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(csr_adaptive_kernel, dim3(nblocks), dim3(kBlock), 0, s,
                       blocks, rp, col, val, x, y);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace clspmv
