// Device helpers shared by the gfx950 kernels: Philox4x32-10, lane-group
// reductions, status plumbing.  Wave64 throughout: a "group" is LPS adjacent
// lanes of one wavefront (LPS a power of two dividing 64), so group
// reductions never cross a wavefront and need no barrier.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace si {

constexpr int kWave = 64;

// Philox4x32-10 (Salmon et al., SC'11): ten rounds of the 4x32 Philox
// S-box with multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments
// 0x9E3779B9 / 0xBB67AE85 between rounds.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k.x += 0x9E3779B9u;
            k.y += 0xBB67AE85u;
        }
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo0 = 0xD2511F53u * c.x;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z);
        const uint32_t lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    }
    return c;
}

// 24-bit uniform in [0, 1), exact in FP32.
__device__ __forceinline__ float u24(uint32_t x) {
    return float(x >> 8) * (1.0f / 16777216.0f);
}

// Butterfly all-reduce over a group of W adjacent lanes.  Addition is
// commutative, so every lane of the group ends with bitwise the same sum.
template <int W>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int o = 1; o < W; o <<= 1) v += __shfl_xor(v, o, W);
    return v;
}

template <int W>
__device__ __forceinline__ float group_max(float v) {
#pragma unroll
    for (int o = 1; o < W; o <<= 1) v = fmaxf(v, __shfl_xor(v, o, W));
    return v;
}

__device__ __forceinline__ uint32_t byte_of(uint32_t w, int i) { return (w >> (8 * i)) & 0xffu; }

}  // namespace si
