// Mixed-precision block SpMM for gfx950: Y += A X with A a BCSR matrix of
// 4x4 blocks whose values are stored in BF16 or FP16, X (k right-hand
// sides) in the same type, and Y in FP32.  The products run on the matrix
// cores: v_mfma_f32_4x4x4 (16 blocks per wave) multiplies, for each group of
// four lanes, one 4x4 block of A by a 4x4 tile of X (four rows of four
// right-hand sides) and accumulates the 4x4 tile of Y in FP32.
//
// Operand layout of the instruction (16 independent 4x4x4 products per
// wave, lane l = 4*g + t for group g):
//   A[i][k]: lane 4g+i, element k      -> row t of the A block (8 bytes)
//   B[k][j]: lane 4g+j, element k      -> X rows c0..c0+3 of column j
//   D[i][j]: lane 4g+j, register i     -> Y rows r0..r0+3 of column j
// so with X and Y column-major every lane moves whole 8- or 16-byte pieces.
// A group owns one (block row, quad of right-hand sides); each Y element
// has one owner, so Y is updated without atomics and the summation order
// (blocks in storage order, four products per instruction) is fixed.

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 h16x4 __attribute__((ext_vector_type(4)));

template <bool BF16>
__device__ __forceinline__ f32x4 mfma4(s16x4 a, s16x4 b, f32x4 c) {
    if constexpr (BF16) {
        return __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a, b, c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_4x4x4f16(__builtin_bit_cast(h16x4, a),
                                                   __builtin_bit_cast(h16x4, b), c, 0, 0, 0);
    }
}

// Loads 4 consecutive 16-bit values at p (8-byte aligned) or, near the end
// of a column, the ones below `lim` and zeros after them.
__device__ __forceinline__ s16x4 ld4(const unsigned short* __restrict__ p, int avail) {
    if (avail >= 4) return *reinterpret_cast<const s16x4*>(p);
    s16x4 v = {0, 0, 0, 0};
    if (avail > 0) v[0] = (short)p[0];
    if (avail > 1) v[1] = (short)p[1];
    if (avail > 2) v[2] = (short)p[2];
    return v;
}

constexpr int kUnroll = 4;

template <bool BF16>
__global__ __launch_bounds__(kBlock) void bspmm4_kernel(
    int brows, int rows, int cols, int k, int nq, const int* __restrict__ rp,
    const int* __restrict__ bcol, const unsigned short* __restrict__ val,
    const unsigned short* __restrict__ X, int ldx, float* __restrict__ Y, int ldy) {
    const int lane = threadIdx.x & (kWave - 1);
    const long long wave = ((long long)blockIdx.x * kBlock + threadIdx.x) / kWave;
    const long long g = wave * 16 + lane / 4;
    const int t = lane & 3;
    const int br = (int)(g / nq), q = (int)(g % nq);
    const bool active = br < brows;
    const int xc = q * 4 + t;  // right-hand side of this lane
    const bool col_ok = active && xc < k;
    int j0 = 0, n = 0;
    if (active) {
        j0 = rp[br];
        n = rp[br + 1] - j0;
    }
    // The instruction is wave-wide: iterate to the longest block row of the
    // wave, feeding zeros past a group's own blocks.
    int nmax = n;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) nmax = max(nmax, __shfl_xor(nmax, m, kWave));
    const unsigned short* xcol = X + (size_t)(col_ok ? xc : 0) * ldx;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int it = 0; it < nmax; it += kUnroll) {
        s16x4 a[kUnroll], b[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            a[u] = s16x4{0, 0, 0, 0};
            b[u] = s16x4{0, 0, 0, 0};
            if (it + u < n) {
                const int j = j0 + it + u;
                a[u] = *reinterpret_cast<const s16x4*>(val + (size_t)j * 16 + t * 4);
                if (col_ok) {
                    const int c0 = bcol[j];
                    b[u] = ld4(xcol + c0, cols - c0);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) acc = mfma4<BF16>(a[u], b[u], acc);
    }
    if (col_ok) {
        const int r0 = br * 4;
        float* yp = Y + (size_t)xc * ldy + r0;
        if (r0 + 3 < rows) {
            f32x4 y = *reinterpret_cast<const f32x4*>(yp);
            y += acc;
            *reinterpret_cast<f32x4*>(yp) = y;
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (r0 + i < rows) yp[i] += acc[i];
        }
    }
}

}  // namespace

namespace launch {

hipError_t bspmm4(hipStream_t s, bool bf16, int brows, int rows, int cols, int k,
                  const int* rp, const int* bcol, const unsigned short* val,
                  const unsigned short* X, int ldx, float* Y, int ldy) {
    if (brows <= 0 || k <= 0) return hipSuccess;
    const int nq = (k + 3) / 4;
    const long long groups = (long long)brows * nq;
    const long long waves = (groups + 15) / 16;
    const long long blocks = (waves * kWave + kBlock - 1) / kBlock;
    if (blocks > 0x7FFFFFFFll) return hipErrorInvalidValue;
    if (bf16)
        hipLaunchKernelGGL(bspmm4_kernel<true>, dim3((unsigned)blocks), dim3(kBlock), 0, s, brows, rows,
                           cols, k, nq, rp, bcol, val, X, ldx, Y, ldy);
    else
        hipLaunchKernelGGL(bspmm4_kernel<false>, dim3((unsigned)blocks), dim3(kBlock), 0, s, brows, rows,
                           cols, k, nq, rp, bcol, val, X, ldx, Y, ldy);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace clspmv
