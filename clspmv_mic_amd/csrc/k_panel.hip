// x-panel SpMV for gfx950: the COO/CSR path for matrices whose columns have
// no locality (power-law rows, random graphs).
//
// Gathering x straight from memory costs a 128-byte L2 line per 4-byte
// element, which caps such kernels near 200 G gathers/s on MI355X whatever
// the HBM bandwidth.  Here the columns are split into panels of at most
// kPanelMax floats; a work-group copies its panel's x into LDS once with
// coalesced loads and then gathers from LDS.  Entries carry a 15-bit
// in-panel column and a segment-start bit in 16 bits, so the matrix stream
// is 6 bytes per nonzero instead of 12 (COO) or 8 (CSR).
//
// Pass 1: each wave takes sub-chunks of 256 entries (4 per lane), forms the
// products, and reduces every segment (one row's run inside the sub-chunk)
// with a segmented scan over the wave's lanes; each segment's sum goes to
// seg_val.  Pass 2: one lane per row with segments adds them in panel order
// and then adds the total to y.  The summation order is fixed by the
// encoding alone; no atomics.

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

constexpr int kU = 2;  // sub-chunks in flight per wave

struct SubChunk {
    u32x2 c;   // four 16-bit entry words
    f32x4 v;   // four values
};

__device__ __forceinline__ void seg_scan_write(const SubChunk& q, int i0, int real_end,
                                               const float* __restrict__ xs, int lane,
                                               float* __restrict__ out) {
    const unsigned w[4] = {q.c.x & 0xFFFFu, q.c.x >> 16, q.c.y & 0xFFFFu, q.c.y >> 16};
    const float v[4] = {q.v.x, q.v.y, q.v.z, q.v.w};
    int f[4];
    float p[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f[k] = (int)(w[k] >> 15);
        p[k] = (i0 + k < real_end) ? v[k] * xs[w[k] & 0x7FFFu] : 0.0f;
    }
    if (lane == 0) f[0] = 1;  // every sub-chunk opens a segment
    // Segment numbering: flags of earlier lanes, then this lane's own.
    const unsigned long long lt = (1ull << lane) - 1ull;
    int before = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) before += __popcll(__ballot(f[k]) & lt);
    const int nf = f[0] + f[1] + f[2] + f[3];

    // Runs inside the lane: head (before the first flag), whole segments,
    // tail (from the last flag on).
    float head = 0.0f, run = 0.0f;
    int id = before - 1;  // segment of the current run
    bool seen = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (f[k]) {
            if (seen) out[id] = run;  // a segment that starts and ends here
            else head = run;
            seen = true;
            ++id;
            run = p[k];
        } else {
            run += p[k];
        }
    }
    if (!seen) head = run;
    // Segmented inclusive scan over lanes of the amount each lane passes on:
    // its tail when it holds a flag, else its whole sum.
    float s = nf ? run : head;
    int g = nf ? 1 : 0;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const float us = __shfl_up(s, d, kWave);
        const int ug = __shfl_up(g, d, kWave);
        if (lane >= d && !g) s += us;
        if (lane >= d) g |= ug;
    }
    // The segment open before this lane ends at its first flag.
    const float prev = __shfl_up(s, 1, kWave);
    if (nf && lane > 0) out[before - 1] = prev + head;
    // The last lane closes the sub-chunk's last segment.
    if (lane == kWave - 1) out[before + nf - 1] = s;
}

template <int NT>
__global__ __launch_bounds__(NT) void panel_pass1_kernel(
    const int4* __restrict__ wg, int pw, int cols,
    const unsigned short* __restrict__ pcol, const float* __restrict__ pval,
    const int* __restrict__ sc_seg, const float* __restrict__ x, bool xa,
    float* __restrict__ seg_val) {
    extern __shared__ float xs[];
    const int4 t = wg[blockIdx.x];
    const int panel = t.x, sc0 = t.y, sc1 = t.z, real_end = t.w;
    const int xb = panel * pw;
    const int w = min(pw, cols - xb);
    const int tid = threadIdx.x;
    if (xa) {
        for (int i = tid * 4; i < w; i += NT * 4) {
            if (i + 3 < w) {
                const float4 v = *reinterpret_cast<const float4*>(x + xb + i);
                xs[i] = v.x;
                xs[i + 1] = v.y;
                xs[i + 2] = v.z;
                xs[i + 3] = v.w;
            } else {
                for (int k = i; k < w; ++k) xs[k] = x[xb + k];
            }
        }
    } else {
        for (int i = tid; i < w; i += NT) xs[i] = x[xb + i];
    }
    __syncthreads();

    const int wave = tid / kWave, lane = tid & (kWave - 1);
    constexpr int nw = NT / kWave;
    for (int sc = sc0 + wave; sc < sc1; sc += nw * kU) {
        SubChunk q[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int c = sc + u * nw;
            if (c < sc1) {
                const int i0 = c * launch::kPanelSub + lane * 4;
                q[u].c = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(pcol + i0));
                q[u].v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(pval + i0));
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int c = sc + u * nw;
            if (c < sc1)
                seg_scan_write(q[u], c * launch::kPanelSub + lane * 4, real_end, xs, lane,
                               seg_val + sc_seg[c]);
        }
    }
}

__global__ __launch_bounds__(kBlock) void panel_pass2_kernel(
    int nrows, const int* __restrict__ rrow, const int* __restrict__ rs_ptr,
    const int* __restrict__ rs_idx, const float* __restrict__ seg_val,
    float* __restrict__ y) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nrows) return;
    const int s = rs_ptr[i], e = rs_ptr[i + 1];
    float acc = 0.0f;
    int j = s;
    // Rows of long power-law tails hold dozens of segments: issue eight
    // independent loads at a time, then add them in order.
    for (; j + 8 <= e; j += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = seg_val[rs_idx[j + k]];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += v[k];
    }
    for (; j < e; ++j) acc += seg_val[rs_idx[j]];
    y[rrow[i]] += acc;
}

}  // namespace

namespace launch {

hipError_t panel_spmv(hipStream_t s, int nwg, const int* wg, int pw, int cols,
                      const unsigned short* pcol, const float* pval,
                      const int* sc_seg, const float* x, bool x_aligned,
                      float* seg_val, int nrows, const int* rrow,
                      const int* rs_ptr, const int* rs_idx, float* y) {
    if (nwg <= 0 || nrows <= 0) return hipSuccess;
    if (pw <= 0 || pw > kPanelMax) return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * (size_t)pw;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void*>(&panel_pass1_kernel<kPanelThreads>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)(sizeof(float) * kPanelMax));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL(panel_pass1_kernel<kPanelThreads>, dim3(nwg), dim3(kPanelThreads), lds, s,
                       reinterpret_cast<const int4*>(wg), pw, cols, pcol, pval, sc_seg, x,
                       x_aligned, seg_val);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(panel_pass2_kernel, dim3((nrows + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       nrows, rrow, rs_ptr, rs_idx, seg_val, y);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace clspmv
