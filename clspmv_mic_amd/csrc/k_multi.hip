// Multi-right-hand-side SpMV for gfx950: Y[:, q] += A X[:, q] for q < k
// (k <= 16) in one pass over the matrix, for DIA and CSR.
//
// X is row-major (the k values of column c are X[c*k .. c*k+k)), so the
// gather of one column brings every right-hand side's value in one or a few
// 16-byte loads: a random column costs one cache line for all k products
// instead of one line per product.  CSR runs over the adaptive row blocks
// (a work-group per block of short rows or per long row).  Y is
// column-major (ldy), so each right-hand side's y stays a contiguous
// vector.  Matrix values may be FP32 or 16-bit (widened on load); sums are
// FP32 in a fixed order, FP64 for CSR rows longer than kLongRowF64; no
// atomics.

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;

// Loads right-hand sides qb .. qb+Q-1 of row c of X (zero where q >= k).
template <int Q>
__device__ __forceinline__ void load_xq(const float* __restrict__ X, int c, int k, int qb, bool vec,
                                        float* xv) {
    const float* p = X + (size_t)c * k + qb;
    if (qb >= k) {  // a lane quad past k (e.g. k = 12 with KB = 16): nothing to load
#pragma unroll
        for (int q = 0; q < Q; ++q) xv[q] = 0.0f;
        return;
    }
    if constexpr (Q == 4) {
        if (vec) {  // k % 4 == 0, X 16-byte aligned
            const float4 t = *reinterpret_cast<const float4*>(p);
            xv[0] = t.x; xv[1] = t.y; xv[2] = t.z; xv[3] = t.w;
            return;
        }
    } else if constexpr (Q == 2) {
        if (vec) {  // k == 2, X 8-byte aligned
            const float2 t = *reinterpret_cast<const float2*>(p);
            xv[0] = t.x; xv[1] = t.y;
            return;
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) xv[q] = qb + q < k ? p[q] : 0.0f;
}

// DIA: L = KB / Q neighbouring lanes share a row, each with Q = min(KB, 4)
// right-hand sides, so a wave's X load for one diagonal is one contiguous
// run of 64 / L rows of X (16-byte pieces).
template <int KB, class V>
__global__ __launch_bounds__(kBlock) void dia_multi_kernel(
    int rows, int cols, int len, int ndiag, const int* __restrict__ off,
    const typename V::T* __restrict__ data, int k, const float* __restrict__ X, bool vec,
    float* __restrict__ Y, int ldy) {
    constexpr int Q = KB < 4 ? KB : 4;
    constexpr int L = KB / Q;
    const int t = threadIdx.x;
    const int qb = (t % L) * Q;
    const int r = xcd_tile(blockIdx.x, gridDim.x) * (kBlock / L) + t / L;
    if (r >= rows) return;
    float acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = 0.0f;
    // C diagonals per step, every load of the step issued before its first
    // product (C * (1 + Q) values in flight per lane); each column is still
    // summed in diagonal order, one product at a time, skipping columns
    // outside x.
    constexpr int C = Q == 4 ? 8 : 16;
    for (int j0 = 0; j0 < ndiag; j0 += C) {  // wave-uniform
        float d[C], xv[C][Q];
        bool in[C];
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int j = min(j0 + u, ndiag - 1);
            d[u] = V::ld(data + (size_t)j * len + r);
            const int c = r + off[j];
            in[u] = j0 + u < ndiag && (unsigned)c < (unsigned)cols;
            load_xq<Q>(X, in[u] ? c : 0, k, qb, vec, xv[u]);
        }
#pragma unroll
        for (int u = 0; u < C; ++u)
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[q] = in[u] ? fmaf(d[u], xv[u][q], acc[q]) : acc[q];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (qb + q < k) Y[(size_t)(qb + q) * ldy + r] += acc[q];
}

// Accumulates a[q] += val[j] * X[col[j]][qb + q] over j = j0, j0 + step, ...
// below j1 in that order, U nonzeros in flight per lane.  Acc is float, or
// double for rows longer than kLongRowF64 (the products are exact in
// double).
template <int Q, class V, class Acc, int U>
__device__ __forceinline__ void csr_multi_run(const int* __restrict__ col,
                                              const typename V::T* __restrict__ val, int j0, int j1,
                                              int step, const float* __restrict__ X, int k, int qb,
                                              bool vec, Acc* a) {
    int j = j0;
    for (; j + (U - 1) * step < j1; j += U * step) {
        int c[U];
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = ld_stream(col + j + u * step);
            v[u] = V::ld(val + j + u * step);
        }
        float xv[U][Q];
#pragma unroll
        for (int u = 0; u < U; ++u) load_xq<Q>(X, c[u], k, qb, vec, xv[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < Q; ++q) a[q] = fma((Acc)v[u], (Acc)xv[u][q], a[q]);
    }
    for (; j < j1; j += step) {
        const int c = ld_stream(col + j);
        const float v = V::ld(val + j);
        float xv[Q];
        load_xq<Q>(X, c, k, qb, vec, xv);
#pragma unroll
        for (int q = 0; q < Q; ++q) a[q] = fma((Acc)v, (Acc)xv[q], a[q]);
    }
}

// One row longer than kAdaptiveNnz by the whole work-group: each lane
// strides the row with 4 * L nonzeros in flight (a whole work-group round
// covers kAdaptiveNnz of them), the four waves' sums are combined in a
// fixed order through LDS, and Y[:, r] is updated once per right-hand side.
template <int KB, class V, class Acc>
__device__ __forceinline__ void csr_multi_long(const typename V::T* __restrict__ val,
                                               const int* __restrict__ col, int r, int j0, int j1,
                                               const float* __restrict__ X, int k, bool vec,
                                               float* __restrict__ Y, int ldy, Acc (*part)[KB]) {
    constexpr int Q = KB < 4 ? KB : 4;
    constexpr int L = KB / Q;
    const int t = threadIdx.x;
    const int qb = (t % L) * Q;
    Acc a[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) a[q] = 0;
    csr_multi_run<Q, V, Acc, launch::kAdaptiveNnz * L / kBlock>(col, val, j0 + t / L, j1, kBlock / L, X, k, qb,
                                                                vec, a);
#pragma unroll
    for (int q = 0; q < Q; ++q)
        for (int m = kWave >> 1; m >= L; m >>= 1) a[q] += __shfl_xor(a[q], m, kWave);
    if ((t & (kWave - 1)) < L) {
#pragma unroll
        for (int q = 0; q < Q; ++q) part[t / kWave][qb + q] = a[q];
    }
    __syncthreads();
    if (t < KB && t < k) {
        Acc tot = 0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) tot += part[w][t];
        float* yp = Y + (size_t)t * ldy + r;
        if constexpr (sizeof(Acc) == sizeof(double))
            *yp = (float)((double)*yp + tot);  // one rounding of the whole row
        else
            *yp += tot;
    }
}

// CSR over the adaptive row blocks (runs of at most kAdaptiveRows rows whose
// nonzeros fit kAdaptiveNnz, or one longer row).  Each lane carries
// Q = min(KB, 4) right-hand sides (one 16-byte gather of X per nonzero), and
// L = KB / Q neighbouring lanes cover the k values of one nonzero.
//
// A block of short rows is streamed: its nonzeros are cut into G = kBlock / L
// equal consecutive chunks, one per lane group, whatever the rows' lengths
// (a row-per-group split left short rows' groups idle beside one long row's
// and put the gathers of a group's rows one after another).  A group loads
// every column, value and X row of its chunk at once, then walks the chunk
// in order: a row that starts and ends inside it is summed there and added
// to Y; the part of a row that began in an earlier chunk goes to LDS (head),
// and a row that continues past the chunk is finished, after a barrier, by
// the group where it starts (its tail, then the later chunks' heads in
// order).  Each row is thus summed in nonzero order with one break per chunk
// boundary, in FP32, the same way every time.
//
// A long row: the whole work-group, the four waves' partial sums combined in
// a fixed order through LDS, in FP64 for rows longer than kLongRowF64
// (products exact, one rounding into y).
template <int KB, class V>
__global__ __launch_bounds__(kBlock) void csr_multi_kernel(
    const int* __restrict__ blocks, const int* __restrict__ rp, const int* __restrict__ col,
    const typename V::T* __restrict__ val, int k, const float* __restrict__ X, bool vec,
    float* __restrict__ Y, int ldy) {
    constexpr int Q = KB < 4 ? KB : 4;
    constexpr int L = KB / Q;
    constexpr int G = kBlock / L;                   // lane groups
    constexpr int CM = launch::kAdaptiveNnz / G;    // nonzeros per group at most
    static_assert(CM * G == launch::kAdaptiveNnz, "a block's nonzeros must split evenly over the groups");
    __shared__ float part[kBlock / kWave][KB];
    __shared__ double part64[kBlock / kWave][KB];
    __shared__ int srp[launch::kAdaptiveRows + 1];  // the block's row starts, from its first nonzero
    __shared__ float head[G][KB];                   // each chunk's part of a row begun earlier
    const int b = blockIdx.x;
    const int r0 = blocks[b], r1 = blocks[b + 1];
    const int nrows = r1 - r0;
    const int t = threadIdx.x;
    const int qb = (t % L) * Q;
    if (nrows == 1 && rp[r1] - rp[r0] > launch::kAdaptiveNnz) {
        if (rp[r1] - rp[r0] > launch::kLongRowF64)
            csr_multi_long<KB, V, double>(val, col, r0, rp[r0], rp[r1], X, k, vec, Y, ldy, part64);
        else
            csr_multi_long<KB, V, float>(val, col, r0, rp[r0], rp[r1], X, k, vec, Y, ldy, part);
        return;
    }
    const int s = rp[r0];
    const int n = rp[r1] - s;  // at most kAdaptiveNnz
    for (int i = t; i <= nrows; i += kBlock) srp[i] = rp[r0 + i] - s;
    __syncthreads();
    if (n == 0) return;  // only empty rows (the whole work-group)
    const int gi = t / L;
    const int C = (n + G - 1) / G;  // <= CM
    const int c0 = min(n, gi * C), c1 = min(n, c0 + C);
    // Every load of the chunk before its first product (slots past the
    // chunk read nonzero s and are never used).
    int cc[CM];
    float vv[CM], xv[CM][Q];
#pragma unroll
    for (int u = 0; u < CM; ++u) {
        const int j = c0 + u < c1 ? s + c0 + u : s;
        cc[u] = ld_stream(col + j);
        vv[u] = V::ld(val + j);
    }
#pragma unroll
    for (int u = 0; u < CM; ++u) load_xq<Q>(X, cc[u], k, qb, vec, xv[u]);
    float acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = 0.0f;
    auto add_y = [&](int r) {
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (qb + q < k) Y[(size_t)(qb + q) * ldy + r0 + r] += acc[q];
    };
    int r = 0, rend = 0;
    bool in_head = false;  // the current row began before this chunk
    if (c0 < c1) {
        // The row holding nonzero c0: the last row starting at or before it
        // (empty rows before it start at the same place and are passed).
        int lo = 0, hi = nrows - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (srp[mid] <= c0) lo = mid;
            else hi = mid - 1;
        }
        r = lo;
        rend = srp[r + 1];
        in_head = srp[r] < c0;
#pragma unroll
        for (int u = 0; u < CM; ++u) {
            const int j = c0 + u;
            if (j < c1) {
                if (j >= rend) {  // row r ended inside the chunk
                    if (in_head) {
#pragma unroll
                        for (int q = 0; q < Q; ++q) head[gi][qb + q] = acc[q];
                    } else {
                        add_y(r);
                    }
                    in_head = false;
                    do {
                        ++r;
                        rend = srp[r + 1];
                    } while (j >= rend);  // empty rows
#pragma unroll
                    for (int q = 0; q < Q; ++q) acc[q] = 0.0f;
                }
#pragma unroll
                for (int q = 0; q < Q; ++q) acc[q] = fmaf(vv[u], xv[u][q], acc[q]);
            }
        }
        if (in_head) {
#pragma unroll
            for (int q = 0; q < Q; ++q) head[gi][qb + q] = acc[q];
        } else if (rend <= c1) {
            add_y(r);
        }
    }
    __syncthreads();
    // The row this chunk ends inside of and began: its tail, then the heads
    // of the later chunks it runs through, in order.
    if (c0 < c1 && !in_head && rend > c1) {
        const int last = (rend - 1) / C;  // the chunk holding the row's last nonzero
        for (int g2 = gi + 1; g2 <= last; ++g2)
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[q] += head[g2][qb + q];
        add_y(r);
    }
}

int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

template <int KB, class V>
hipError_t dia_multi_kb(hipStream_t s, int rows, int cols, int len, int ndiag, const int* off,
                        const void* data, int k, const float* X, float* Y, int ldy) {
    constexpr int Q = KB < 4 ? KB : 4;
    const bool vec = Q > 1 && k % Q == 0 && ((uintptr_t)X % (Q * sizeof(float))) == 0;
    hipLaunchKernelGGL((dia_multi_kernel<KB, V>), dim3(ceil_div(rows, kBlock / (KB / Q))), dim3(kBlock), 0,
                       s, rows, cols, len, ndiag, off, static_cast<const typename V::T*>(data), k, X, vec, Y,
                       ldy);
    return hipGetLastError();
}

template <class V>
hipError_t dia_multi_v(hipStream_t s, int rows, int cols, int len, int ndiag, const int* off,
                       const void* data, int k, const float* X, float* Y, int ldy) {
    if (k <= 1) return dia_multi_kb<1, V>(s, rows, cols, len, ndiag, off, data, k, X, Y, ldy);
    if (k <= 2) return dia_multi_kb<2, V>(s, rows, cols, len, ndiag, off, data, k, X, Y, ldy);
    if (k <= 4) return dia_multi_kb<4, V>(s, rows, cols, len, ndiag, off, data, k, X, Y, ldy);
    if (k <= 8) return dia_multi_kb<8, V>(s, rows, cols, len, ndiag, off, data, k, X, Y, ldy);
    return dia_multi_kb<16, V>(s, rows, cols, len, ndiag, off, data, k, X, Y, ldy);
}

template <int KB, class V>
hipError_t csr_multi_kb(hipStream_t s, int nblocks, const int* blocks, const int* rp, const int* col,
                        const void* val, int k, const float* X, float* Y, int ldy) {
    constexpr int Q = KB < 4 ? KB : 4;
    const bool vec = Q > 1 && k % Q == 0 && ((uintptr_t)X % (Q * sizeof(float))) == 0;
    hipLaunchKernelGGL((csr_multi_kernel<KB, V>), dim3(nblocks), dim3(kBlock), 0, s, blocks, rp, col,
                       static_cast<const typename V::T*>(val), k, X, vec, Y, ldy);
    return hipGetLastError();
}

template <class V>
hipError_t csr_multi_v(hipStream_t s, int nblocks, const int* blocks, const int* rp, const int* col,
                       const void* val, int k, const float* X, float* Y, int ldy, int kb) {
    if (kb > 0) {  // a fixed width: the summation order of each right-hand side does not depend on k
        switch (kb) {
            case 1: return csr_multi_kb<1, V>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy);
            case 2: return csr_multi_kb<2, V>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy);
            case 4: return csr_multi_kb<4, V>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy);
            case 8: return csr_multi_kb<8, V>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy);
            case 16: return csr_multi_kb<16, V>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy);
            default: return hipErrorInvalidValue;
        }
    }
    if (k <= 1) return csr_multi_kb<1, V>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy);
    if (k <= 2) return csr_multi_kb<2, V>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy);
    if (k <= 4) return csr_multi_kb<4, V>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy);
    if (k <= 8) return csr_multi_kb<8, V>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy);
    return csr_multi_kb<16, V>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy);
}

}  // namespace

namespace launch {

hipError_t dia_multi(hipStream_t s, int rows, int cols, int len, int ndiag, const int* offsets,
                     int dtype, const void* data, int k, const float* X, float* Y, int ldy) {
    if (rows <= 0 || ndiag <= 0 || k <= 0) return hipSuccess;
    if (k > kMultiMax) return hipErrorInvalidValue;
    switch (dtype) {
        case 0: return dia_multi_v<ValF32>(s, rows, cols, len, ndiag, offsets, data, k, X, Y, ldy);
        case 1: return dia_multi_v<ValBF16>(s, rows, cols, len, ndiag, offsets, data, k, X, Y, ldy);
        case 2: return dia_multi_v<ValF16>(s, rows, cols, len, ndiag, offsets, data, k, X, Y, ldy);
        default: return hipErrorInvalidValue;
    }
}

hipError_t csr_multi(hipStream_t s, int nblocks, const int* blocks, const int* rp, const int* col,
                     int dtype, const void* val, int k, const float* X, float* Y, int ldy, int kb) {
    if (nblocks <= 0 || k <= 0) return hipSuccess;
    if (k > kMultiMax || (kb > 0 && k > kb)) return hipErrorInvalidValue;
    switch (dtype) {
        case 0: return csr_multi_v<ValF32>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy, kb);
        case 1: return csr_multi_v<ValBF16>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy, kb);
        case 2: return csr_multi_v<ValF16>(s, nblocks, blocks, rp, col, val, k, X, Y, ldy, kb);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace launch
}  // namespace clspmv
