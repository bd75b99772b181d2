// PBWT index build and per-marker state table for gfx950.
//
// PBWT (reference: PBWTBuilder::build_marker, src/pbwt/pbwt_index.cpp:153-208):
// the order after marker m is a stable partition of the order after m-1 by the
// allele at m (allele == 0 first, every nonzero allele after), and every
// divergence in row m equals m (p and q are reset to m after each push,
// pbwt_index.cpp:173-192).
//
// State table (reference: PBWTIndex::select_states with all keys equal,
// pbwt_index.cpp:20-44, pairing imputer.cpp:913-920, allele reads
// emission.cu:96-103): states[m][k] = prefix[m][perm[k]] for k < H, k % H
// beyond, partner (h + 1) % H.  Sample-independent, so it is built once.
#include "si_device.hpp"
#include "si_internal.hpp"

namespace si {

// ---------------------------------------------------------------------------
// Serial sweep, one workgroup: thread t owns the contiguous positions
// [t*K, t*K + K) of the previous order.  Per marker: gather the allele bit of
// each owned haplotype from the column staged in LDS, count zeros, workgroup
// exclusive scan of the counts, scatter into the new row.  The previous row is
// the prefix row written one step before (global memory; the workgroup's
// waves share one CU and its write-through L1, so the barrier orders it).
// ---------------------------------------------------------------------------
constexpr int kSweepThreads = 1024;
constexpr int kSweepMaxK = 64;  // H <= 65536

__global__ __launch_bounds__(kSweepThreads) void pbwt_sweep_serial_kernel(
    const uint64_t* __restrict__ cols, uint32_t H, uint32_t M, uint32_t W,
    uint32_t* __restrict__ prefix, uint32_t* __restrict__ div) {
    extern __shared__ uint64_t col_lds[];  // [W]
    __shared__ uint32_t wave_tot[kSweepThreads / kWave];
    __shared__ uint32_t zero_total;

    const uint32_t t = threadIdx.x;
    const uint32_t K = (H + kSweepThreads - 1) / kSweepThreads;
    const uint32_t i0 = t * K;
    const uint32_t lane = t & (kWave - 1), wid = t / kWave;

    uint32_t hap[kSweepMaxK];
    for (uint32_t m = 0; m < M; ++m) {
        for (uint32_t w = t; w < W; w += kSweepThreads) col_lds[w] = cols[size_t(m) * W + w];
        __syncthreads();

        const uint32_t* prev = (m == 0) ? nullptr : prefix + size_t(m - 1) * H;
        uint32_t zeros = 0;
        uint64_t xbits = 0;  // bit j set iff allele of owned position j is nonzero
#pragma unroll 4
        for (uint32_t j = 0; j < K; ++j) {
            const uint32_t i = i0 + j;
            uint32_t h = 0;
            if (i < H) {
                h = prev ? prev[i] : i;
                const uint32_t x = uint32_t(col_lds[h >> 6] >> (h & 63)) & 1u;
                xbits |= uint64_t(x) << j;
                zeros += x ^ 1u;
            }
            hap[j] = h;
        }
        // workgroup exclusive scan of zero counts
        uint32_t incl = zeros;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t n = __shfl_up(incl, o, kWave);
            if (lane >= uint32_t(o)) incl += n;
        }
        if (lane == kWave - 1) wave_tot[wid] = incl;
        __syncthreads();
        if (t == 0) {
            uint32_t run = 0;
            for (uint32_t w = 0; w < kSweepThreads / kWave; ++w) {
                const uint32_t v = wave_tot[w];
                wave_tot[w] = run;
                run += v;
            }
            zero_total = run;
        }
        __syncthreads();
        const uint32_t Z = zero_total;
        uint32_t zb = wave_tot[wid] + incl - zeros;  // zeros before position i0
        uint32_t* row = prefix + size_t(m) * H;
#pragma unroll 4
        for (uint32_t j = 0; j < K; ++j) {
            const uint32_t i = i0 + j;
            if (i < H) {
                const bool one = (xbits >> j) & 1u;
                const uint32_t pos = one ? Z + (i - zb) : zb;
                row[pos] = hap[j];
                zb += one ? 0u : 1u;
            }
        }
        if (div) {
            uint32_t* drow = div + size_t(m) * H;
            for (uint32_t i = t; i < H; i += kSweepThreads) drow[i] = m;
        }
        __syncthreads();
    }
}

int launch_pbwt_sweep(const uint64_t* cols, uint32_t H, uint32_t M, uint32_t* prefix,
                      uint32_t* div, hipStream_t stream) {
    if (H == 0 || M == 0) return 0;
    if (H > uint32_t(kSweepThreads) * kSweepMaxK) return SI_ERR_UNSUPPORTED;
    const uint32_t W = (H + 63) / 64;
    const size_t lds = size_t(W) * sizeof(uint64_t);
    if (lds > 64 * 1024) return SI_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(pbwt_sweep_serial_kernel, dim3(1), dim3(kSweepThreads), lds, stream, cols,
                       H, M, W, prefix, div);
    return hipGetLastError() == hipSuccess ? 0 : SI_ERR_DEVICE;
}

// ---------------------------------------------------------------------------
// State table: one thread per marker.
// ---------------------------------------------------------------------------
struct PermArg {
    uint32_t v[64];
};

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
    uint32_t present = 0;
    for (uint32_t k = 0; k < Lp; ++k) {
        uint8_t b;
        if (k < L) {
            uint32_t h0, h1;
            if (mode == 1) {
                h0 = 2 * k;
                h1 = 2 * k + 1;
            } else {
                h0 = (k < H) ? prefix[size_t(m) * H + perm.v[k]] : (k % H);
                h1 = (h0 + 1 == H) ? 0u : h0 + 1;
            }
            if (states) states[size_t(m) * L + k] = h0;
            const uint32_t a0 = allele_raw(cols, miss, W, m, h0);
            const uint32_t a1 = allele_raw(cols, miss, W, m, h1);
            const uint32_t sum = a0 + a1;
            const uint32_t cls = sum >= 2 ? 2u : sum;
            b = uint8_t((a0 == 1) | ((a1 == 1) << 1) | ((a0 == 255) << 2) | ((a1 == 255) << 3) |
                        (cls << 4));
            present |= 1u << cls;
        } else {
            b = uint8_t(0x40 | (3u << 4));  // padding state
        }
        sa[size_t(m) * Lp + k] = b;
    }
    mmask[m] = uint8_t(present);
}

int launch_state_table(const uint32_t* prefix, const uint64_t* cols, const uint64_t* miss,
                       uint32_t H, uint32_t M, uint32_t L, uint32_t Lp, uint32_t mode,
                       const uint32_t* perm_host, uint32_t* states, uint8_t* sa, uint8_t* mmask,
                       hipStream_t stream) {
    if (M == 0) return 0;
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
