// Device-side helpers shared by the gfx950 SpMV kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clspmv {
namespace dev {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlock = 256;        // work-group size of every row kernel
constexpr int kXcds = 8;           // MI355X: 8 XCDs, each with its own L2

// Matrix arrays are read exactly once per SpMV: load them non-temporally so
// they do not evict the x vector from L2 / MALL.
template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

// 128-bit non-temporal loads (the address must be 16-byte aligned).
__device__ __forceinline__ float4 ld_stream(const float4* p) {
    const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 ld_stream(const int4* p) {
    const i32x4 v = __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(p));
    return make_int4(v.x, v.y, v.z, v.w);
}

// Stored matrix values: FP32, or BF16 / FP16 bits widened to FP32 when
// loaded (products and sums stay FP32).  ld: one value; ld2 / ld4: two /
// four consecutive values (8 / 16-byte aligned for FP32, 4 / 8 for 16-bit).
struct ValF32 {
    typedef float T;
    static __device__ __forceinline__ float ld(const float* p) { return ld_stream(p); }
    static __device__ __forceinline__ float2 ld2(const float* p) {
        return *reinterpret_cast<const float2*>(p);
    }
    static __device__ __forceinline__ float4 ld4(const float* p) {
        return ld_stream(reinterpret_cast<const float4*>(p));
    }
    static __device__ __forceinline__ float4 ld4c(const float* p) {  // cached load
        return *reinterpret_cast<const float4*>(p);
    }
};

template <class Cvt>
struct Val16 {
    typedef unsigned short T;
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ float ld(const T* p) { return Cvt::f(ld_stream(p)); }
    static __device__ __forceinline__ float2 ld2(const T* p) {
        const unsigned w = ld_stream(reinterpret_cast<const unsigned*>(p));
        return make_float2(Cvt::f((T)(w & 0xFFFFu)), Cvt::f((T)(w >> 16)));
    }
    static __device__ __forceinline__ float4 ld4(const T* p) {
        return unpack(__builtin_nontemporal_load(reinterpret_cast<const u32x2*>(p)));
    }
    static __device__ __forceinline__ float4 ld4c(const T* p) {  // cached load
        return unpack(*reinterpret_cast<const u32x2*>(p));
    }
    static __device__ __forceinline__ float4 unpack(u32x2 w) {
        return make_float4(Cvt::f((T)(w.x & 0xFFFFu)), Cvt::f((T)(w.x >> 16)),
                           Cvt::f((T)(w.y & 0xFFFFu)), Cvt::f((T)(w.y >> 16)));
    }
};
struct CvtBF16 {
    static __device__ __forceinline__ float f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
};
struct CvtF16 {
    static __device__ __forceinline__ float f(unsigned short h) {
        return (float)__builtin_bit_cast(_Float16, h);
    }
};
typedef Val16<CvtBF16> ValBF16;
typedef Val16<CvtF16> ValF16;

// Work-groups are dispatched round-robin over the XCDs (group b runs on XCD
// b % 8).  Remap so that XCD k processes one contiguous range of tiles: rows
// that share x windows (banded matrices) then share one L2.  Bijective on
// [0, nb) for any nb.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int q = nb / kXcds, r = nb % kXcds;
    const int x = b % kXcds, i = b / kXcds;
    return x * q + (x < r ? x : r) + i;
}

// The lane's value from another lane of its 16-lane row, by DPP (no LDS
// crossbar round trip): CTRL 0xB1 = quad_perm(1,0,3,2) (lane ^ 1), 0x4E =
// quad_perm(2,3,0,1) (lane ^ 2), 0x141 = row_half_mirror (lane i <- 7 - i
// within each 8 lanes), 0x140 = row_mirror (lane i <- 15 - i).
template <int CTRL>
__device__ __forceinline__ float dpp_row(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// The value of lane ^ M (M a power of two below 64): DPP quad permutes for
// 1 and 2, a ds_swizzle in bit-mask mode (and-mask 0x1F, xor-mask M: no
// address computation, unlike ds_bpermute) within each 32 lanes for 4..16,
// a shuffle across the wave's halves for 32.
template <int M>
__device__ __forceinline__ float xor_lane(float v) {
    static_assert(M >= 1 && M <= 32 && (M & (M - 1)) == 0, "lane xor mask");
    if constexpr (M == 1) return dpp_row<0xB1>(v);
    else if constexpr (M == 2) return dpp_row<0x4E>(v);
    else if constexpr (M < 32)
        return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1F | (M << 10)));
    else return __shfl_xor(v, M, kWave);
}

// Sum over aligned groups of G consecutive lanes (G a power of two, at most
// 64); every lane of a group receives the group's total (the same bits in
// every lane: each step adds a pair of values held by both partners).
// Steps inside a 16-lane row use DPP: after lane ^ 1 and lane ^ 2 each lane
// holds its quad's sum, the half mirror pairs the two quads of 8 lanes and
// the row mirror the two halves of 16; larger groups finish with xor_lane.
template <int G>
__device__ __forceinline__ float group_sum(float v) {
    if constexpr (G >= 2) v += dpp_row<0xB1>(v);
    if constexpr (G >= 4) v += dpp_row<0x4E>(v);
    if constexpr (G >= 8) v += dpp_row<0x141>(v);
    if constexpr (G >= 16) v += dpp_row<0x140>(v);
    if constexpr (G >= 32) v += xor_lane<16>(v);
    if constexpr (G >= 64) v += xor_lane<32>(v);
    return v;
}

// Sum over the lanes of an aligned group of g lanes (a power of two, at
// most 64, uniform across the wave) that share lane % l: v + v(lane ^ m)
// for m = g/2 down to l, every lane of a set receiving the same bits.
__device__ __forceinline__ float xor_sum(float v, int g, int l) {
    if (g >= 64 && l <= 32) v += xor_lane<32>(v);
    if (g >= 32 && l <= 16) v += xor_lane<16>(v);
    if (g >= 16 && l <= 8) v += xor_lane<8>(v);
    if (g >= 8 && l <= 4) v += xor_lane<4>(v);
    if (g >= 4 && l <= 2) v += xor_lane<2>(v);
    if (g >= 2 && l <= 1) v += xor_lane<1>(v);
    return v;
}

// The whole group's sum with a run-time group size.
__device__ __forceinline__ float group_sum_rt(float v, int g) { return xor_sum(v, g, 1); }

// x[e..e+3] as a float4.  The fast path needs e+3 < n and 16-byte alignment
// of x (checked once on the host, `x_aligned`); otherwise each element is
// loaded separately and elements at or beyond n read as zero.
__device__ __forceinline__ float4 ld_x4(const float* __restrict__ x, int e, int n,
                                        bool x_aligned) {
    if (x_aligned && e + 3 < n) return *reinterpret_cast<const float4*>(x + e);
    float4 v;
    v.x = (e + 0 < n) ? x[e + 0] : 0.0f;
    v.y = (e + 1 < n) ? x[e + 1] : 0.0f;
    v.z = (e + 2 < n) ? x[e + 2] : 0.0f;
    v.w = (e + 3 < n) ? x[e + 3] : 0.0f;
    return v;
}

__device__ __forceinline__ float dot4(float4 a, float4 b, float acc) {
    acc = fmaf(a.x, b.x, acc);
    acc = fmaf(a.y, b.y, acc);
    acc = fmaf(a.z, b.z, acc);
    acc = fmaf(a.w, b.w, acc);
    return acc;
}

}  // namespace dev
}  // namespace clspmv
