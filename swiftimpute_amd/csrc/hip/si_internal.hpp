// Internal declarations shared by the HIP translation units of libsi_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../../include/swiftimpute_hip.h"

namespace si {

// Smallest checkpoint block of the forward-backward passes (markers).
constexpr uint32_t kMinCkptBlock = 16;
// Largest state count L (padded to 256: 32 or 64 lanes of 8 or 4 states).
constexpr uint32_t kMaxStates = 256;
// Largest genotype-likelihood palette the kernels stage in LDS.
constexpr uint32_t kMaxPalette = 64;

// Constants of the untyped step (every real state's emission 1) under the
// job's constant transitions, computed on the host in FP64: S = L t_o + u,
// c = u / S.  ok = 0 disables untyped steps (per-marker transitions, or a
// shape outside 0 < S < 1; log form: t_o < 2^-30).
struct URun {
    float c1, c2;        // u / S, t_o / S
    float rS, rL;        // 1 / S, 1 / L
    float lc1, lnc1;     // log2 c, ln c
    float lnS, gS;       // ln S, 1 / (1 - S)
    uint32_t ok;
    double lnSd;         // ln S
};

// One job as a lane group sees it (plain and batched forms alike).
struct JobView {
    uint32_t B, M, L, flags, pitch, rng_stream, sample_base, nblk;  // nblk: checkpoint blocks
    uint32_t W, nwin;        // window length and count (W = 0: one window, no restarts)
    uint64_t seed;
    float4 step;             // (t_o, u, lto, lu) when not per-marker
    const float4* steps;     // [M] per-marker coefficients
    const uint8_t* sa;       // [M][Lp]
    const uint8_t* mmask;    // [M]
    const uint8_t* codes;    // [B][pitch]
    float* ds;
    float* ap;
    uint8_t* hap;
    float* prob;
    double* loglik;
    float* ck_alpha;  // [B][nblk][Lp]  forward vector at each block's first marker
    float* ck_beta;   // [B][nblk][Lp]  backward vector at each block's last marker
    double* ck_bl;    // [B][nblk]      its log scale
    float* ck_win;    // [B][nwin][Lp]  forward vector at each window's last marker
    uint8_t* ck_aband;  // [B][nblk][Lp] bands of ck_alpha (banded jobs)
    uint8_t* ck_bband;  // [B][nblk][Lp] bands of ck_beta (banded jobs)
    uint32_t* tmask;  // [ceil(M/16)]   typed markers of each load block (fast jobs)
    uint32_t* tlist;  // [M]            the typed markers 1..M-1 in order (fast jobs)
    uint32_t* tcount; // [1]            how many
    URun ur;
};

// Emission tables shared by every job of a launch.
struct FbTables {
    const float* eps;      // [P][8][4] linear kernel
    const float* emax;     // [P][8]
    const float* palette;  // [P][3]   log-space kernel
    uint32_t P;
    uint32_t ucode;        // code of an untyped marker (kNoUntyped: none)
};
constexpr uint32_t kNoUntyped = 0xffffffffu;
// The untyped code a job asks for (SI_FB_UNTYPED), or kNoUntyped.
inline uint32_t fb_untyped_code(const si_fb_job& j) {
    return ((j.flags & SI_FB_UNTYPED) && j.untyped_code < j.n_palette) ? j.untyped_code : kNoUntyped;
}

// Forward-backward kernel variant: lanes per sample in the chain and block
// passes (states per lane = Lp / lanes) and the checkpoint block length.
struct FbVariant {
    uint32_t gc, gb, cb;
};
constexpr uint32_t kMaxCkptBlock = 32;
FbVariant fb_default_variant(uint32_t Lp);
// Lane counts instantiated for a pass (0: chain, 1: block); returns how many.
uint32_t fb_candidates(uint32_t Lp, int pass, uint32_t* lanes, uint32_t max);
bool fb_variant_valid(uint32_t Lp, const FbVariant& v);

// Workspace of one job at the smallest checkpoint block (enough for all).
// Window length a job runs with: 0 unless windowing applies (window > 0 and
// M > window, src/api/imputer.cpp:965-966).
inline uint32_t fb_window(const si_fb_job& j) {
    return (j.window > 0 && j.n_markers > j.window) ? j.window : 0u;
}
constexpr uint32_t kMinWindow = 32;
// Banded jobs (SI_FB_BANDED on a log-space job without windows) run the
// banded kernels and keep the bands of their checkpoints in the workspace.
inline bool fb_band_job(const si_fb_job& j) {
    return (j.flags & SI_FB_BANDED) && (j.flags & SI_FB_LOGSPACE) && fb_window(j) == 0u;
}
size_t fb_workspace_bytes_one(uint32_t B, uint32_t M, uint32_t Lp, uint32_t W, bool band);
void fb_bind_workspace(JobView& v, uint32_t Lp, char* ws);
JobView fb_job_view(const si_fb_job& j, uint32_t cb);

// Plain form: one job, one launch per pass.
int launch_fb(const JobView& v, const FbTables& t, uint32_t Lp, uint32_t passes,
              const FbVariant& var, hipStream_t stream);
// Load blocks of a job's typed-marker mask per workgroup of its pass.
constexpr uint32_t kTmaskBlocksPerGroup = 32;
__host__ __device__ inline uint64_t fb_tmask_groups(uint32_t M) {
    return (uint64_t((M + 15) / 16) + kTmaskBlocksPerGroup - 1) / kTmaskBlocksPerGroup;
}

// A batched class of fast jobs (untyped steps): class flag beside SI_FB_LOGSPACE.
constexpr uint32_t kFbFastClass = 1u << 31;
// Whether a job runs the untyped-step kernels (with 16-marker checkpoint
// blocks): constant transitions the closed-form runs accept, no windows, an
// untyped code in the palette (the kernels still check its likelihoods).
URun fb_urun(const si_fb_job& j);
inline bool fb_fast_job(const si_fb_job& j) {
    return fb_untyped_code(j) < 256u && fb_window(j) == 0u && fb_urun(j).ok != 0u && !fb_band_job(j);
}

// Batched form: samples of a job padded to a multiple of 64 in both queues, so
// a wavefront's round of groups never spans two jobs (its job descriptor and
// untyped-marker mask are wavefront-uniform).
constexpr uint32_t kBatchSamplePad = 64;
__host__ __device__ inline uint64_t fb_batch_padded_samples(uint32_t B) {
    return (uint64_t(B) + kBatchSamplePad - 1) / kBatchSamplePad * kBatchSamplePad;
}

// Batched form: every job of one padded-state class in one persistent launch
// per pass.  jobs: device array in launch order; item_off: device prefix of
// per-job item counts (n_jobs + 1) for the chain pass (2 chains per sample)
// and for the block pass (one item per (sample, checkpoint block)); counters:
// 2 device 64-bit words, zeroed by the launcher.
struct BatchClass {
    uint32_t Lp;
    uint32_t n_jobs;
    uint32_t flags;  // SI_FB_LOGSPACE, SI_FB_BANDED, kFbFastClass
    const JobView* jobs;
    const uint64_t* chain_off;
    const uint64_t* block_off;
    const uint64_t* mask_off;  // workgroups of the typed-marker pass per job
    uint64_t chain_items, block_items, mask_groups;
    unsigned long long* counters;  // [2] queue positions (chain, block)
    uint32_t grid_chain, grid_block;  // persistent grid sizes (workgroups)
};
// Persistent grid sizes for a variant (as many workgroups as fit at once,
// capped by the work).
void fb_batch_grid(BatchClass& c, const FbVariant& var, uint32_t P);
int launch_fb_batch(const BatchClass& c, const FbTables& t, uint32_t passes, const FbVariant& var,
                    hipStream_t stream);

// Per-(Lp, lanes) kernel sets, registered by the instantiation units.
enum FbForm : int { kFbPlainChain = 0, kFbPlainBlock = 1, kFbBatchChain = 2, kFbBatchBlock = 3 };
struct FbLaunch {
    int form;
    bool log;
    uint32_t cb;
    bool fast;             // untyped-step kernels (16-marker checkpoint blocks)
    bool band;             // banded log-space kernels
    const JobView* job;    // plain forms
    uint32_t Bpad;         // plain chain: forward groups padded to whole wavefronts
    const BatchClass* bc;  // batched forms
    FbTables tb;
    uint32_t grid;
    hipStream_t stream;
};
struct FbImpl {
    uint32_t Lp, G;
    bool has_block;
    int (*launch)(const FbLaunch&);
    int (*occupancy)(int form, bool log, uint32_t cb, uint32_t P, bool fast, bool band);  // workgroups per CU
};
void fb_register(const FbImpl& impl);
const FbImpl* fb_impl(uint32_t Lp, uint32_t G);

// PBWT prefix/divergence construction (pbwt.hip): chunks of C markers.
constexpr uint32_t kPbwtMaxH = 1u << 19;  // the global-memory form's two LDS column buffers: 128 KiB
uint32_t pbwt_default_chunk(uint32_t H, uint32_t M);
size_t pbwt_workspace_bytes(uint32_t H, uint32_t M, uint32_t C);
int launch_pbwt_sweep(const uint64_t* cols, uint32_t H, uint32_t M, uint32_t* prefix,
                      uint32_t* div, uint32_t C, void* ws, size_t ws_bytes, hipStream_t stream);
int launch_state_table(const uint32_t* prefix, const uint64_t* cols, const uint64_t* miss,
                       uint32_t H, uint32_t M, uint32_t L, uint32_t Lp, uint32_t mode,
                       const uint32_t* perm_host, uint32_t* states, uint8_t* sa, uint8_t* mmask,
                       hipStream_t stream);

}  // namespace si
