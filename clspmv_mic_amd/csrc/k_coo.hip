// COO SpMV for gfx950 (replaces kernels/spmv_coo.cl gpu_coo_s1/gpu_coo_s2).
//
// Input: entries sorted by row.  Each work-group takes one chunk of
// kCooChunk consecutive entries, 4 per lane, loaded as 128-bit vectors.
//   1. A lane sums runs of equal rows among its 4 products.  A run that both
//      starts and ends inside the lane is a whole row and goes straight to y.
//   2. A segmented inclusive scan over the lanes' trailing runs (wave
//      shuffles, then one LDS exchange across the 4 waves) gives each lane
//      the part of its leading run that earlier lanes hold.
//   3. Rows completed inside the chunk are added to y.  The chunk's first
//      and last rows may continue into neighbouring chunks: their partial sums
//      go to two carry slots per chunk instead.
// A second launch walks the carry slots (ordered by chunk, so equal rows are
// adjacent) and adds each row's slots, in chunk order, to y.  No atomics: the
// summation order is fixed by the chunking alone.
//
// The reference's s1 kernel drops the last 32 padded entries when the
// aligned nnz is an odd multiple of 32 (kernels/spmv_coo.cl:15); this kernel
// covers every entry.

#include <climits>

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;
constexpr int kV = launch::kCooChunk / kBlock;  // entries per lane (4)
static_assert(kV == 4, "vector loads assume 4 entries per lane");

struct Seg {
    int flag;    // 1: a row starts inside (or at the start of) this lane
    float val;
};

__device__ __forceinline__ Seg seg_combine(Seg a, Seg b) {  // a precedes b
    return Seg{a.flag | b.flag, b.flag ? b.val : a.val + b.val};
}

__global__ __launch_bounds__(kBlock) void coo_chunk_kernel(
    int nnz, const int* __restrict__ row, const int* __restrict__ col,
    const float* __restrict__ val, const float* __restrict__ x,
    float* __restrict__ y, int* __restrict__ carry_row,
    float* __restrict__ carry_val) {
    __shared__ int head_rows[kBlock];
    __shared__ int tail_rows[kBlock];
    __shared__ Seg wave_tot[kBlock / kWave];

    const int c = blockIdx.x;
    const int t = threadIdx.x;
    const int wave = t / kWave, lane = t & (kWave - 1);
    const int base = c * launch::kCooChunk;
    const int last = min(base + launch::kCooChunk, nnz) - 1;
    const int chunk_head = row[base];
    const int chunk_tail = row[last];

    // ---- load 4 entries; entries past the end extend the last row with 0.
    int r[kV];
    float p[kV];
    const int i0 = base + t * kV;
    if (i0 + kV - 1 <= last) {
        const int4 r4 = ld_stream(reinterpret_cast<const int4*>(row + i0));
        const int4 c4 = ld_stream(reinterpret_cast<const int4*>(col + i0));
        const float4 v4 = ld_stream(reinterpret_cast<const float4*>(val + i0));
        r[0] = r4.x; r[1] = r4.y; r[2] = r4.z; r[3] = r4.w;
        p[0] = v4.x * x[c4.x];
        p[1] = v4.y * x[c4.y];
        p[2] = v4.z * x[c4.z];
        p[3] = v4.w * x[c4.w];
    } else {
#pragma unroll
        for (int v = 0; v < kV; ++v) {
            const int i = i0 + v;
            if (i <= last) {
                r[v] = row[i];
                p[v] = val[i] * x[col[i]];
            } else {
                r[v] = chunk_tail;
                p[v] = 0.0f;
            }
        }
    }

    // ---- runs inside the lane
    const int hr = r[0];
    float head_sum = p[0], run = p[0];
    int cur = r[0];
    bool split = false;   // lane holds at least one row boundary
#pragma unroll
    for (int v = 1; v < kV; ++v) {
        if (r[v] == cur) {
            run += p[v];
        } else {
            if (!split) head_sum = run;
            else y[cur] += run;  // a whole row inside this lane
            split = true;
            cur = r[v];
            run = p[v];
        }
    }
    const int tr = cur;
    const float tail_sum = run;
    if (!split) head_sum = run;

    head_rows[t] = hr;
    tail_rows[t] = tr;
    __syncthreads();
    const int prev_tail = t > 0 ? tail_rows[t - 1] : INT_MIN;
    const int next_head = t + 1 < kBlock ? head_rows[t + 1] : INT_MIN;

    // ---- segmented inclusive scan of trailing runs over the lanes
    Seg s{(split || t == 0 || hr != prev_tail) ? 1 : 0, tail_sum};
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const float uv = __shfl_up(s.val, d, kWave);
        const int uf = __shfl_up(s.flag, d, kWave);
        if (lane >= d) s = seg_combine(Seg{uf, uv}, s);
    }
    if (lane == kWave - 1) wave_tot[wave] = s;
    __syncthreads();
    Seg pre{0, 0.0f};  // combined totals of earlier waves
    for (int w = 0; w < wave; ++w) pre = seg_combine(pre, wave_tot[w]);
    const Seg incl = wave > 0 ? seg_combine(pre, s) : s;
    // Inclusive value of lane t-1.
    float prev_incl = __shfl_up(incl.val, 1, kWave);
    if (lane == 0) prev_incl = pre.val;

    // ---- rows completed in this lane
    if (split) {
        // The leading run ends inside this lane.
        const float tot = head_sum + ((t > 0 && prev_tail == hr) ? prev_incl : 0.0f);
        if (hr == chunk_head) {
            carry_row[2 * c] = hr;
            carry_val[2 * c] = tot;
        } else {
            y[hr] += tot;
        }
    }
    if (t + 1 < kBlock) {
        // The trailing run ends at this lane's last entry.
        if (next_head != tr) {
            if (tr == chunk_head) {
                carry_row[2 * c] = tr;
                carry_val[2 * c] = incl.val;
            } else {
                y[tr] += incl.val;
            }
        }
    } else {
        carry_row[2 * c + 1] = tr;  // tr == chunk_tail
        carry_val[2 * c + 1] = incl.val;
        if (chunk_head == chunk_tail) {
            // One row fills the chunk: the head slot joins the same row.
            carry_row[2 * c] = chunk_head;
            carry_val[2 * c] = 0.0f;
        }
    }
}

__global__ __launch_bounds__(kBlock) void coo_carry_kernel(
    int nslots, const int* __restrict__ carry_row,
    const float* __restrict__ carry_val, float* __restrict__ y) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nslots) return;
    const int r = carry_row[i];
    if (i > 0 && carry_row[i - 1] == r) return;  // not the first slot of r
    float tot = 0.0f;
    for (int j = i; j < nslots && carry_row[j] == r; ++j) tot += carry_val[j];
    y[r] += tot;
}

}  // namespace

namespace launch {

hipError_t coo_segreduce(hipStream_t s, int nnz, const int* row, const int* col,
                         const float* val, const float* x, float* y,
                         int* carry_row, float* carry_val) {
    if (nnz <= 0) return hipSuccess;
    const int chunks = (nnz + kCooChunk - 1) / kCooChunk;
    hipLaunchKernelGGL(coo_chunk_kernel, dim3(chunks), dim3(kBlock), 0, s, nnz,
                       row, col, val, x, y, carry_row, carry_val);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int slots = 2 * chunks;
    hipLaunchKernelGGL(coo_carry_kernel, dim3((slots + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, slots, carry_row, carry_val, y);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace clspmv
