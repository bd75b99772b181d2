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

// DPP lane exchanges within a row of 16 lanes: quad_perm [1,0,3,2] (xor 1),
// quad_perm [2,3,0,1] (xor 2), row_half_mirror (i <-> 7-i within 8) and
// row_mirror (i <-> 15-i within 16).
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// All-reduce sum over a group of W adjacent lanes (W = 1, 2, 4, 8, 16, aligned).
// After the xor steps every lane of a quad holds the quad sum; the mirror
// steps then pair whole quads / half-rows.  Addition is commutative, so every
// lane of the group ends with bitwise the same sum.
template <int W>
__device__ __forceinline__ float group_sum(float v) {
    static_assert(W == 1 || W == 2 || W == 4 || W == 8 || W == 16, "group width");
    if constexpr (W >= 2) v += dpp<0xB1>(v);
    if constexpr (W >= 4) v += dpp<0x4E>(v);
    if constexpr (W >= 8) v += dpp<0x141>(v);
    if constexpr (W >= 16) v += dpp<0x140>(v);
    return v;
}

template <int W>
__device__ __forceinline__ float group_max(float v) {
    static_assert(W == 1 || W == 2 || W == 4 || W == 8 || W == 16, "group width");
    if constexpr (W >= 2) v = fmaxf(v, dpp<0xB1>(v));
    if constexpr (W >= 4) v = fmaxf(v, dpp<0x4E>(v));
    if constexpr (W >= 8) v = fmaxf(v, dpp<0x141>(v));
    if constexpr (W >= 16) v = fmaxf(v, dpp<0x140>(v));
    return v;
}

// Natural log and exp on the hardware log2/exp2 (arguments are normal
// numbers here, so no denormal scaling is needed).
__device__ __forceinline__ float fast_ln(float x) {
    return __builtin_amdgcn_logf(x) * 0.69314718055994531f;
}
__device__ __forceinline__ float fast_exp(float x) {
    return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
}

__device__ __forceinline__ uint32_t byte_of(uint32_t w, int i) { return (w >> (8 * i)) & 0xffu; }

}  // namespace si
