// Internal declarations shared by the HIP translation units of libsi_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../../include/swiftimpute_hip.h"

namespace si {

// Markers per forward-backward block: the checkpoint interval of the chain
// pass and the replay length of the block pass.
constexpr uint32_t kBlock = 16;
// Largest genotype-likelihood palette the kernels stage in LDS.
constexpr uint32_t kMaxPalette = 64;

// Kernel-side view of one job (passed by value).
struct FbParams {
    uint32_t B, M, L, Lp, flags, pitch, P, rng_stream, sample_base;
    uint32_t nblk;  // ceil(M / kBlock)
    uint32_t Bpad;  // forward-chain groups, rounded so no wavefront mixes directions
    uint64_t seed;
    si_step_coef step;
    const si_step_coef* steps;
    const uint8_t* sa;
    const uint8_t* mmask;
    const uint8_t* codes;
    const float* eps;
    const float* emax;
    const float* palette;
    float* ds;
    float* ap;
    uint8_t* hap;
    float* prob;
    double* loglik;
    float* ck_alpha;  // [B][nblk][Lp]  forward vector at each block's first marker
    float* ck_beta;   // [B][nblk][Lp]  backward vector at each block's last marker
    double* ck_bl;    // [B][nblk]      its log scale
};

size_t fb_workspace_bytes_one(uint32_t B, uint32_t M, uint32_t Lp);
int launch_fb(const FbParams& p, uint32_t passes, hipStream_t stream);

int launch_pbwt_sweep(const uint64_t* cols, uint32_t H, uint32_t M, uint32_t* prefix,
                      uint32_t* div, hipStream_t stream);
int launch_state_table(const uint32_t* prefix, const uint64_t* cols, const uint64_t* miss,
                       uint32_t H, uint32_t M, uint32_t L, uint32_t Lp, uint32_t mode,
                       const uint32_t* perm_host, uint32_t* states, uint8_t* sa, uint8_t* mmask,
                       hipStream_t stream);

}  // namespace si
