// Counter-based uniform values shared by host and device: x_j[i] of a
// request or report is a pure function of (seed, j, i), so a host can
// regenerate exactly what a kernel used.
#pragma once

#include <hip/hip_runtime.h>

namespace clspmv {

// SplitMix64 finalizer: a bijective 64-bit mix.
__host__ __device__ inline unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// 24 random bits scaled to [0, 2) minus 1: exact in FP32, so host and
// device agree bit for bit.
__host__ __device__ inline float uniform_at(unsigned long long seed, int j, int i) {
    const unsigned long long h =
        mix64(seed ^ mix64(((unsigned long long)(unsigned)j << 32) | (unsigned)i));
    return (float)(h >> 40) * (1.0f / 8388608.0f) - 1.0f;
}

// FP32 -> BF16 / FP16 bits, round to nearest even (finite inputs).
__device__ inline unsigned short to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ inline unsigned short to_f16(float f) {
    const _Float16 h = (_Float16)f;
    return __builtin_bit_cast(unsigned short, h);
}

}  // namespace clspmv
