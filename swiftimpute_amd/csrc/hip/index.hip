// Per-marker state table for gfx950 (the PBWT sweep itself is in pbwt.hip).
//
// State table (reference: PBWTIndex::select_states with all keys equal,
// pbwt_index.cpp:20-44, pairing imputer.cpp:913-920, allele reads
// emission.cu:96-103): states[m][k] = prefix[m][perm[k]] for k < H, k % H
// beyond, partner (h + 1) % H; a perm entry flagged SI_PERM_FIXED names the
// haplotype itself.  Sample-independent, so it is built once.
#include "si_device.hpp"
#include "si_internal.hpp"

namespace si {

// ---------------------------------------------------------------------------
// State table: one thread per marker.
// ---------------------------------------------------------------------------
struct PermArg {
    uint32_t v[kMaxStates];  // a kernel argument: every selected slot's prefix position
};
static_assert(sizeof(PermArg) <= 1024, "state permutation fits the kernel arguments (4 KiB)");

__device__ __forceinline__ uint32_t allele_raw(const uint64_t* cols, const uint64_t* miss,
                                               uint32_t W, uint32_t m, uint32_t h) {
    const size_t w = size_t(m) * W + (h >> 6);
    if (miss && ((miss[w] >> (h & 63)) & 1ull)) return 255u;
    return uint32_t(cols[w] >> (h & 63)) & 1u;
}

__global__ __launch_bounds__(256) void state_table_kernel(
    const uint32_t* __restrict__ prefix, const uint64_t* __restrict__ cols,
    const uint64_t* __restrict__ miss, uint32_t H, uint32_t M, uint32_t L, uint32_t Lp,
    uint32_t mode, PermArg perm, uint32_t* __restrict__ states, uint8_t* __restrict__ sa,
    uint8_t* __restrict__ mmask) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const uint32_t W = (H + 63) / 64;
    // state pair of slot k and its allele byte (without the mask bit)
    auto slot = [&](uint32_t k, uint32_t& cls) -> uint8_t {
        uint32_t h0, h1;
        if (mode == 1) {
            h0 = 2 * k;
            h1 = 2 * k + 1;
        } else {
            const uint32_t e = (k < H) ? perm.v[k] : (SI_PERM_FIXED | (k % H));
            h0 = (e & SI_PERM_FIXED) ? (e & ~SI_PERM_FIXED) : prefix[size_t(m) * H + e];
            h1 = (h0 + 1 == H) ? 0u : h0 + 1;
        }
        if (states) states[size_t(m) * L + k] = h0;
        const uint32_t a0 = allele_raw(cols, miss, W, m, h0);
        const uint32_t a1 = allele_raw(cols, miss, W, m, h1);
        const uint32_t sum = a0 + a1;
        cls = sum >= 2 ? 2u : sum;
        // the class sits at bits 2-3, so a byte-select and yields class * 4, the
        // state's byte offset into an emission row
        return uint8_t((a0 == 1) | ((a1 == 1) << 1) | (cls << 2) | ((a0 == 255) << 4) |
                       ((a1 == 255) << 5));
    };
    uint32_t present = 0;
    for (uint32_t k = 0; k < L; ++k) {
        uint32_t cls;
        (void)slot(k, cls);
        present |= 1u << cls;
    }
    for (uint32_t k = 0; k < Lp; ++k) {
        uint32_t cls = 3;
        uint8_t b = (k < L) ? slot(k, cls) : uint8_t(0x40 | (3u << 2));  // padding state
        // every lane's 4-byte word carries the marker's class-present mask in
        // bit 7 of its bytes 0-2
        if ((k & 3u) < 3u) b |= uint8_t(((present >> (k & 3u)) & 1u) << 7);
        sa[size_t(m) * Lp + k] = b;
    }
    mmask[m] = uint8_t(present);
}

int launch_state_table(const uint32_t* prefix, const uint64_t* cols, const uint64_t* miss,
                       uint32_t H, uint32_t M, uint32_t L, uint32_t Lp, uint32_t mode,
                       const uint32_t* perm_host, uint32_t* states, uint8_t* sa, uint8_t* mmask,
                       hipStream_t stream) {
    if (M == 0) return 0;
    if (L > kMaxStates) return SI_ERR_ARG;  // the permutation argument holds kMaxStates slots
    PermArg p{};
    if (mode == 0) {
        const uint32_t n = L < H ? L : H;
        for (uint32_t k = 0; k < n; ++k) p.v[k] = perm_host[k];
    }
    const uint32_t threads = 256;
    hipLaunchKernelGGL(state_table_kernel, dim3((M + threads - 1) / threads), dim3(threads), 0,
                       stream, prefix, cols, miss, H, M, L, Lp, mode, p, states, sa, mmask);
    return hipGetLastError() == hipSuccess ? 0 : SI_ERR_DEVICE;
}

}  // namespace si
