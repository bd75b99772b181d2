// One lane shape of the forward-backward kernels: compiled once per
// (SI_FB_LP padded states, SI_FB_G lanes per sample), each object registering
// its launchers with the host side (hmm.hip) when the library loads.  Block
// kernels exist for R = Lp / G <= 4; R = 8 is a chain-pass shape only.
#include "fb_kernels.hpp"

#if !defined(SI_FB_LP) || !defined(SI_FB_G)
#error "compile with -DSI_FB_LP=<padded states> -DSI_FB_G=<lanes per sample>"
#endif

namespace si {
namespace {

constexpr int kLp = SI_FB_LP, kG = SI_FB_G, kR = kLp / kG;
constexpr bool kHasBlock = kR <= 4;

// Chain pass of fast jobs (untyped steps, 16-marker checkpoint blocks only);
// their block pass is the exact kernel's.
template <bool LOG>
int launch_fast_chain(const FbLaunch& L) {
    const size_t dyn = fb_table_lds_bytes(L.tb.P);
    const dim3 grid(L.grid), block(kFbThreads);
    if (L.form == kFbPlainChain)
        hipLaunchKernelGGL((fb_chain_fast_kernel<kG, kR, LOG, 16>), grid, block, dyn, L.stream, *L.job, L.tb,
                           L.Bpad);
    else
        hipLaunchKernelGGL((fb_chain_fast_batch_kernel<kG, kR, LOG, 16>), grid, block, dyn, L.stream, *L.bc,
                           L.tb);
    return hipGetLastError() == hipSuccess ? 0 : SI_ERR_DEVICE;
}

template <bool LOG, int CB>
int launch_t(const FbLaunch& L) {
    const size_t dyn = fb_table_lds_bytes(L.tb.P);
    const dim3 grid(L.grid), block(kFbThreads);
    switch (L.form) {
        case kFbPlainChain:
            hipLaunchKernelGGL((fb_chain_kernel<kG, kR, LOG, CB>), grid, block, dyn, L.stream, *L.job,
                               L.tb, L.Bpad);
            break;
        case kFbBatchChain:
            hipLaunchKernelGGL((fb_chain_batch_kernel<kG, kR, LOG, CB>), grid, block, dyn, L.stream,
                               *L.bc, L.tb);
            break;
        case kFbPlainBlock:
            if constexpr (kHasBlock) {
                hipLaunchKernelGGL((fb_block_kernel<kG, kR, LOG, CB>), grid, block, dyn, L.stream,
                                   *L.job, L.tb);
                break;
            } else {
                return SI_ERR_UNSUPPORTED;
            }
        case kFbBatchBlock:
            if constexpr (kHasBlock) {
                hipLaunchKernelGGL((fb_block_batch_kernel<kG, kR, LOG, CB>), grid, block, dyn, L.stream,
                                   *L.bc, L.tb);
                break;
            } else {
                return SI_ERR_UNSUPPORTED;
            }
        default:
            return SI_ERR_ARG;
    }
    return hipGetLastError() == hipSuccess ? 0 : SI_ERR_DEVICE;
}

// Banded log-space kernels (SI_FB_BANDED jobs, 16-marker checkpoint blocks).
int launch_band(const FbLaunch& L) {
    const size_t dyn = fb_table_lds_bytes_band(L.tb.P);
    const dim3 grid(L.grid), block(kFbThreads);
    switch (L.form) {
        case kFbPlainChain:
            hipLaunchKernelGGL((fb_chain_band_kernel<kG, kR>), grid, block, dyn, L.stream, *L.job, L.tb, L.Bpad);
            break;
        case kFbBatchChain:
            hipLaunchKernelGGL((fb_chain_band_batch_kernel<kG, kR>), grid, block, dyn, L.stream, *L.bc, L.tb);
            break;
        case kFbPlainBlock:
            if constexpr (kHasBlock) {
                hipLaunchKernelGGL((fb_block_band_kernel<kG, kR>), grid, block, dyn, L.stream, *L.job, L.tb);
                break;
            } else {
                return SI_ERR_UNSUPPORTED;
            }
        case kFbBatchBlock:
            if constexpr (kHasBlock) {
                hipLaunchKernelGGL((fb_block_band_batch_kernel<kG, kR>), grid, block, dyn, L.stream, *L.bc,
                                   L.tb);
                break;
            } else {
                return SI_ERR_UNSUPPORTED;
            }
        default:
            return SI_ERR_ARG;
    }
    return hipGetLastError() == hipSuccess ? 0 : SI_ERR_DEVICE;
}

int launch(const FbLaunch& L) {
    if (L.band) return (L.cb == 16 && L.log) ? launch_band(L) : SI_ERR_UNSUPPORTED;
    if (L.fast && (L.form == kFbPlainChain || L.form == kFbBatchChain))
        return L.cb == 16 ? (L.log ? launch_fast_chain<true>(L) : launch_fast_chain<false>(L))
                          : SI_ERR_UNSUPPORTED;
    if (L.cb == 16) return L.log ? launch_t<true, 16>(L) : launch_t<false, 16>(L);
    return SI_ERR_UNSUPPORTED;
}

template <bool LOG>
int occupancy_fast_chain(int form, uint32_t P) {
    int n = 0;
    const size_t dyn = fb_table_lds_bytes(P);
    const hipError_t e =
        form == kFbPlainChain
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fb_chain_fast_kernel<kG, kR, LOG, 16>, kFbThreads, dyn)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fb_chain_fast_batch_kernel<kG, kR, LOG, 16>,
                                                           kFbThreads, dyn);
    return (e == hipSuccess && n > 0) ? n : 1;
}

template <bool LOG, int CB>
int occupancy_t(int form, uint32_t P) {
    int n = 0;
    const size_t dyn = fb_table_lds_bytes(P);
    hipError_t e = hipErrorInvalidValue;
    switch (form) {
        case kFbPlainChain:
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fb_chain_kernel<kG, kR, LOG, CB>,
                                                             kFbThreads, dyn);
            break;
        case kFbBatchChain:
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, fb_chain_batch_kernel<kG, kR, LOG, CB>, kFbThreads, dyn);
            break;
        case kFbPlainBlock:
            if constexpr (kHasBlock)
                e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fb_block_kernel<kG, kR, LOG, CB>,
                                                                 kFbThreads, dyn);
            break;
        case kFbBatchBlock:
            if constexpr (kHasBlock)
                e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &n, fb_block_batch_kernel<kG, kR, LOG, CB>, kFbThreads, dyn);
            break;
        default:
            break;
    }
    return (e == hipSuccess && n > 0) ? n : 1;
}

int occupancy_band(int form, uint32_t P) {
    int n = 0;
    const size_t dyn = fb_table_lds_bytes_band(P);
    hipError_t e = hipErrorInvalidValue;
    switch (form) {
        case kFbPlainChain:
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fb_chain_band_kernel<kG, kR>, kFbThreads, dyn);
            break;
        case kFbBatchChain:
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fb_chain_band_batch_kernel<kG, kR>, kFbThreads,
                                                             dyn);
            break;
        case kFbPlainBlock:
            if constexpr (kHasBlock)
                e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fb_block_band_kernel<kG, kR>, kFbThreads, dyn);
            break;
        case kFbBatchBlock:
            if constexpr (kHasBlock)
                e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fb_block_band_batch_kernel<kG, kR>, kFbThreads,
                                                                 dyn);
            break;
        default:
            break;
    }
    return (e == hipSuccess && n > 0) ? n : 1;
}

int occupancy(int form, bool log, uint32_t cb, uint32_t P, bool fast, bool band) {
    if (band) return occupancy_band(form, P);
    if (fast && (form == kFbPlainChain || form == kFbBatchChain))
        return log ? occupancy_fast_chain<true>(form, P) : occupancy_fast_chain<false>(form, P);
    return log ? occupancy_t<true, 16>(form, P) : occupancy_t<false, 16>(form, P);
}

const bool registered = (fb_register(FbImpl{kLp, kG, kHasBlock, launch, occupancy}), true);

}  // namespace
}  // namespace si
