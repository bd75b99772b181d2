/*
 * swiftimpute_hip.h -- the C ABI of the MI355X (gfx950) imputation hot path.
 *
 * Plain C: device pointers, sizes, an opaque stream handle (a hipStream_t
 * passed as void*, NULL = the null stream).  No allocation happens inside an
 * entry point that takes a workspace; every such call is asynchronous on the
 * given stream and returns a status (0 = SI_OK, negative = error, message in
 * si_last_error(), which is thread-local).
 *
 * The C++ API in include/swiftimpute/imputer.hpp (Imputer, ReferencePanel,
 * TargetData, ImputationConfig, ImputationResult) is implemented on top of
 * these entry points.  Each declaration names the reference interface it
 * replaces (paths relative to the tcBio/SwiftImpute tree).
 *
 * Data layouts (all little-endian, row-major):
 *   cols     [M][W] uint64, W = ceil(H/64): bit (h & 63) of word (h >> 6) of
 *            row m is 1 iff the panel allele at (m, h) is nonzero.
 *   miss     [M][W] uint64 or NULL: bit set iff the allele is missing (255).
 *   prefix   [M][H] uint32 -- PBWT order after marker m.
 *   div      [M][H] uint32 -- PBWT divergence (every entry of row m is m).
 *   sa       [M][Lp] uint8  -- state-allele table, Lp = si_padded_states(L):
 *              bit0 allele(h0)==1, bit1 allele(h1)==1, bits2-3 emission
 *              class min(a0+a1,2) (3 for padding), bit4 h0 missing, bit5 h1
 *              missing, bit6 padding state (k >= L), bit7 of byte k (k % 4 < 3) =
 *              bit (k % 4) of the marker's mmask (so every 4-byte word of
 *              the row carries the mask).
 *   mmask    [M] uint8      -- bit c set iff some real state has class c.
 *   codes    [B][pitch] uint8 -- genotype-likelihood palette index per
 *            (sample, marker); pitch >= M, multiple of 16.
 *   palette  [P][3] float   -- log10 genotype likelihoods (ll_00, ll_01, ll_11)
 *            (GenotypeLikelihoods, src/core/types.hpp:46-52).
 *   eps      [P][8][4] float -- emission table built by si_emission_table.
 *   emax     [P][8] float   -- its log offsets.
 *   ds       [B][M] float   -- posterior dosage, ap [B][M][2] float,
 *   hap      [B][2][M] uint8 (ImputationResult layout, src/api/imputer.hpp:191),
 *   prob     [B][M][L] float -- the sampler's normalized state probabilities.
 */
#ifndef SWIFTIMPUTE_HIP_H
#define SWIFTIMPUTE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SI_ABI_VERSION 1

enum si_status {
    SI_OK = 0,
    SI_ERR_ARG = -1,         /* invalid argument (reference: ImputationError) */
    SI_ERR_DEVICE = -2,      /* HIP runtime error (reference: CUDAError)      */
    SI_ERR_UNSUPPORTED = -3, /* shape or mode outside what this build handles */
    SI_ERR_WORKSPACE = -4    /* workspace missing or too small                */
};

/* Thread-local message of the last failing call. */
const char* si_last_error(void);
int si_abi_version(void);

/* Padded state count the kernels use: the next of 4, 8, 16, ..., 256 >= L. */
uint32_t si_padded_states(uint32_t L);

/* ------------------------------------------------------------------ PBWT --
 * Replaces PBWTBuilder::build / build_parallel / build_marker
 * (src/pbwt/pbwt_index.cpp:95-247, declared src/pbwt/pbwt_index.hpp:105-135),
 * which the reference runs on the host.
 *   cols   [M][ceil(H/64)] device, bit h of word h/64 set iff allele != 0
 *   prefix [M][H] uint32 device; div [M][H] uint32 device or NULL (not written)
 * H <= 524288 (2^19).  The sweep works in chunks of markers: by default, for
 * H <= 32768, c = ceil(M / (7/8 of the current device's compute units)),
 * rounded up to a multiple of 128 when above 64 and at least 16 (256 units
 * assumed when no device answers), and 512 above.  The
 * workspace (si_pbwt_workspace_bytes, device; query it with the device of the
 * sweep current) may be NULL, in which case the call takes stream-ordered
 * scratch of its own.  Asynchronous on `stream`.
 */
size_t si_pbwt_workspace_bytes(uint32_t H, uint32_t M);
int si_pbwt_sweep(const uint64_t* cols, uint32_t H, uint32_t M, uint32_t* prefix,
                  uint32_t* div, void* workspace, size_t workspace_bytes, void* stream);
/* The same with an explicit chunk length (markers per chunk, >= 1). */
size_t si_pbwt_chunked_workspace_bytes(uint32_t H, uint32_t M, uint32_t chunk);
int si_pbwt_sweep_chunked(const uint64_t* cols, uint32_t H, uint32_t M, uint32_t* prefix,
                          uint32_t* div, uint32_t chunk, void* workspace, size_t workspace_bytes,
                          void* stream);

/* ------------------------------------------------------------- selection --
 * Host: the permutation libstdc++ std::sort applies to H equal keys under the
 * reference comparator (PBWTIndex::select_states, src/pbwt/pbwt_index.cpp:20-44).
 * perm receives min(L, H) entries.
 */
int si_select_perm(uint32_t H, uint32_t L, uint32_t* perm);

/* A perm entry with this bit set names a haplotype directly (the low 31 bits)
 * instead of a position in the marker's prefix row. */
#define SI_PERM_FIXED 0x80000000u

/* Host: the selection the reference's GPU kernels make instead
 * (launch_select_states, src/pbwt/pbwt_selector.cu:400-463), which differs from
 * the CPU path above.  With every divergence of a row equal:
 *   L <= 32 and H <= 10000 (select_states_warp_kernel, :212-312): prefix
 *     positions 0, 32, 64, 96 (as many as lane 0 holds, min(4, ceil(H/32))),
 *     then haplotypes 4, 5, ..., flagged SI_PERM_FIXED;
 *   otherwise, L <= 64 (select_states_bitonic_kernel, :320-394) and L > 64
 *     (select_states_kernel, :47-204): positions 0..L-1.
 * For H < 128 the reference's warp kernel shuffles across lanes that have
 * left its loop (undefined); this function takes lane 0's entries.  Where a
 * thread's range of the reference's L > 64 kernel holds fewer than L
 * haplotypes, its merge also reads slots that thread never wrote (undefined);
 * this function takes them as empty.  perm receives min(L, H) entries;
 * L <= 256. */
int si_select_perm_gpu_compat(uint32_t H, uint32_t L, uint32_t* perm);

/* Device: L-best state selection + pairing + allele lookup, replacing the
 * reference's per-(sample, marker) CPU selection and pair expansion
 * (src/api/imputer.cpp:894-951) and the emission kernels' allele reads
 * (src/kernels/emission.cu:96-103).  The selection is sample-independent, so
 * one [M][L] table serves every sample.
 *   mode 0: PBWT selection, states[m][k] = prefix[m][perm[k]] (k < H), k % H
 *           otherwise; partner (h + 1) % H.  An entry flagged SI_PERM_FIXED
 *           gives the haplotype itself (perm[k] & 0x7fffffff).  Entries must
 *           name a position or haplotype below H.
 *   mode 1: simple selection (use_pbwt_selection = false), pair (2k, 2k+1)
 *           (src/api/imputer.cpp:933-942); requires 2L <= H.
 * perm_host: min(L, H) entries from si_select_perm or
 *   si_select_perm_gpu_compat (ignored in mode 1).
 * states [M][L] uint32 (h0) may be NULL.  sa/mmask as described above.
 */
int si_state_table(const uint32_t* prefix, const uint64_t* cols, const uint64_t* miss,
                   uint32_t H, uint32_t M, uint32_t L, uint32_t mode,
                   const uint32_t* perm_host, uint32_t* states, uint8_t* sa, uint8_t* mmask,
                   void* stream);

/* ----------------------------------------------------------- transitions --
 * Host: Li-Stephens constants per interval t = 0..M-2 from the genetic map in
 * Morgans (TransitionComputer::compute, src/kernels/transition.cu:83-121,
 * 242-289, with the map of src/api/imputer.cpp:596-605).  FP64 math, FP32
 * results: Td (stay) and To (switch), log10 values as the reference stores.
 */
int si_genetic_map(const uint64_t* pos, const double* cM, uint32_t M, double* g);
int si_transition_constants(const double* g, uint32_t M, uint32_t ne, float rho_rate,
                            float* Td, float* To);

/* Per-step coefficients the kernels consume for one interval:
 *   t_o = e^To, u = e^Td - e^To (linear space), lto = To, lu = log(e^Td - e^To).
 * The reference combines its log10 constants with natural exp/log
 * (src/kernels/logsumexp.cuh:67-70); these coefficients keep that. */
typedef struct si_step_coef {
    float t_o;
    float u;
    float lto;
    float lu;
} si_step_coef;
int si_step_coefs(const float* Td, const float* To, uint32_t n, si_step_coef* out);

/* ------------------------------------------------------------- emission --
 * Host: per palette entry p and class-present mask c (1..7), the emission
 * factors eps[p][c][k] = exp(palette[p][k] - emax[p][c]) for each class k
 * present, where emax[p][c] = max over present classes (FP64 exp, FP32 out).
 * Restates compute_single_emission (src/kernels/emission.cu:35-54) in the
 * max-shifted linear form the forward-backward consumes.
 */
int si_emission_table(const float* palette, uint32_t P, float* eps, float* emax);

/* -------------------------------------------------------- forward-backward --
 * One imputation job: B samples over M markers with L states.  Replaces
 * EmissionComputer::compute, launch_forward_pass, launch_backward_pass and
 * HaplotypeSampler::sample (src/kernels/emission.cu:208-257,
 * src/kernels/forward_backward.cu:321-388, src/kernels/sampling.cu:269-320),
 * fused; adds the posterior dosage the reference leaves unfilled
 * (ImputationResult::set_posterior, src/api/imputer.cpp:459-465).
 */
enum si_fb_flags {
    SI_FB_PER_MARKER = 1u << 0, /* steps[] per marker; else `step` for every
                                   marker, as the reference (interval 0 only,
                                   forward_backward.cu:108)                   */
    SI_FB_ARGMAX = 1u << 1,     /* argmax instead of the Philox draw           */
    SI_FB_LOGSPACE = 1u << 2,   /* log-space FP32 kernel instead of the
                                   rescaled linear-space one                   */
    SI_FB_UNTYPED = 1u << 3     /* markers whose code is `untyped_code` (a
                                   palette entry with all three likelihoods
                                   0: a missing genotype, every state's
                                   emission 1) take the untyped step: the
                                   same affine map for every sample, composed
                                   over runs of such markers in O(1) per
                                   marker instead of O(L).  Same model, FP32
                                   rounding only; ignored when that palette
                                   entry is not (0, 0, 0)                     */,
    SI_FB_BANDED = 1u << 4      /* log-space jobs whose switch log-probability
                                   can fall to the -999 floor (coincident
                                   markers): every value carried as a residual
                                   beside a count of -999 floors, so values on
                                   the floor keep FP32 precision relative to
                                   it (posteriors within 1e-5 of FP64 where
                                   the plain log form gives 5e-5).  Ignored
                                   for linear and windowed jobs              */
};

typedef struct si_fb_job {
    uint32_t n_samples;   /* B */
    uint32_t n_markers;   /* M */
    uint32_t n_states;    /* L (1..256) */
    uint32_t flags;       /* si_fb_flags */
    uint32_t code_pitch;  /* row pitch of codes, >= M, multiple of 16 */
    uint32_t n_palette;   /* P (1..64) */
    uint32_t rng_stream;  /* third Philox counter word */
    uint32_t sample_base; /* global index of sample 0 (Philox counter word 1) */
    uint64_t seed;        /* Philox key */
    si_step_coef step;    /* used when SI_FB_PER_MARKER is clear */
    const si_step_coef* steps; /* [M] device, entry m = step into marker m */
    const uint8_t* sa;    /* [M][Lp] device */
    const uint8_t* mmask; /* [M] device */
    const uint8_t* codes; /* [B][code_pitch] device */
    const float* eps;     /* [P][8][4] device (linear kernel) */
    const float* emax;    /* [P][8] device (both kernels: the FP64-carried offset) */
    const float* palette; /* [P][3] device (log-space kernel) */
    float* ds;            /* [B][M] device or NULL */
    float* ap;            /* [B][M][2] device or NULL */
    uint8_t* hap;         /* [B][2][M] device or NULL */
    float* prob;          /* [B][M][L] device or NULL */
    double* loglik;       /* [B] device or NULL */
    uint32_t window;      /* 0, or the reference's windowed model when M > window
                             (>= 32): run_windowed_forward_backward,
                             src/api/imputer.cpp:1044-1174.  Window w > 0 starts
                             its forward from the previous window's last vector
                             (no transition or emission into marker w*window),
                             its posteriors use a forward re-run from the
                             window's own last vector, and the backward step out
                             of each window start is not taken
                             (forward_backward.cu:406-626).  loglik is that of
                             the first forward run. */
    uint32_t untyped_code; /* with SI_FB_UNTYPED: the palette code of an
                              untyped marker (< 256)                          */
} si_fb_job;

/* Workspace (device bytes) si_fb_run needs for these jobs (enough for every
 * kernel variant). */
size_t si_fb_workspace_bytes(const si_fb_job* jobs, uint32_t n_jobs);

/* Kernel variant: lanes of a wavefront per sample in the chain pass and in
 * the block pass (states per lane = Lp / lanes, Lp = si_padded_states(L)),
 * and the checkpoint block length in markers (16; other lengths are refused
 * with SI_ERR_UNSUPPORTED).  Every sum over states is the same binary tree
 * whatever the variant, so outputs are bit-identical across variants: the
 * choice is performance only.  No counterpart in the reference (one thread
 * per (sample, state), forward_backward.cu:41-317). */
typedef struct si_fb_variant {
    uint32_t chain_lanes;
    uint32_t block_lanes;
    uint32_t ckpt_block;
} si_fb_variant;

/* The variant si_fb_run uses for L states (four states per lane, 16-marker
 * blocks). */
int si_fb_default_variant(uint32_t n_states, si_fb_variant* out);

/* Lane counts built for a pass (pass = SI_FB_PASS_CHAIN or SI_FB_PASS_BLOCK)
 * at L states, written to lanes[0..max); returns how many exist. */
uint32_t si_fb_lane_candidates(uint32_t n_states, uint32_t pass, uint32_t* lanes, uint32_t max);

/* Plain form: one launch sequence per job, jobs processed in order on the
 * stream.  jobs is a host array. */
int si_fb_run(const si_fb_job* jobs, uint32_t n_jobs, void* workspace, size_t workspace_bytes,
              void* stream);

/* The two launches si_fb_run issues per job, selectable: SI_FB_PASS_CHAIN
 * (forward and backward chains writing block checkpoints into the workspace)
 * and SI_FB_PASS_BLOCK (block replay, posteriors, dosage, sampling, reading
 * them).  Running CHAIN then BLOCK on the same workspace equals si_fb_run;
 * callers use the split to time the passes or to pipeline batches across
 * streams.  No counterpart in the reference (its forward and backward are
 * separate launches, forward_backward.cu:321-388). */
enum si_fb_pass { SI_FB_PASS_CHAIN = 1u, SI_FB_PASS_BLOCK = 2u };
int si_fb_run_passes(const si_fb_job* jobs, uint32_t n_jobs, void* workspace,
                     size_t workspace_bytes, uint32_t passes, void* stream);

/* si_fb_run_passes with an explicit variant (NULL: the default). */
int si_fb_run_variant(const si_fb_job* jobs, uint32_t n_jobs, const si_fb_variant* variant,
                      void* workspace, size_t workspace_bytes, uint32_t passes, void* stream);

/* Allocation-free form of si_fb_run_variant (both passes), for callers that
 * cannot allocate device memory: code under stream capture, tight serving
 * loops, harnesses that forbid hipMalloc.  No counterpart in the reference.
 * The checkpoint workspace is a static device array of
 * SI_FB_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load until the
 * process ends; jobs whose si_fb_workspace_bytes exceed it are refused with
 * SI_ERR_WORKSPACE.  The call allocates nothing and runs on the current
 * device.  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (other streams, other host threads) race, so serialize them.
 * Nothing in the array needs to start cleared.  variant may be NULL (the
 * default).  Outputs are identical to si_fb_run's. */
#define SI_FB_NOALLOC_SCRATCH_BYTES (64u << 20)
int si_fb_run_noalloc(const si_fb_job* jobs, uint32_t n_jobs, const si_fb_variant* variant,
                      void* stream);

/* Autotuner for the plain form: times every candidate variant on these jobs
 * (one padded-state class, one numeric space), each job cut to its first
 * max_markers markers (0: 2048) with every output disabled, so the jobs'
 * output buffers are not touched; writes the fastest to *out.  The workspace
 * is the one si_fb_run takes for these jobs.  Synchronizes the stream. */
int si_fb_tune(const si_fb_job* jobs, uint32_t n_jobs, void* workspace, size_t workspace_bytes,
               uint32_t max_markers, void* stream, si_fb_variant* out);

/* ---------------------------------------------------------- batched form --
 * Many jobs of uneven cohort size, marker count and L in one persistent
 * work-queue launch per pass and padded-state class (jobs longest first),
 * optionally replayed from a captured hipGraph.  No counterpart in the
 * reference, which runs one fixed-size batch after another
 * (src/api/imputer.cpp:765-777).
 *
 * All jobs share one emission table (eps/emax/palette pointers and
 * n_palette of jobs[0]; the others must match).  si_fb_batch_create plans the
 * batch and allocates its device tables (set-up, not a hot call);
 * si_fb_batch_run copies the tables on its stream, from page-locked staging,
 * ahead of the launches, and waits on the host only for the previous run's
 * table copies.  The workspace holds every job's checkpoints at once.
 */
typedef struct si_fb_batch si_fb_batch;
int si_fb_batch_create(const si_fb_job* jobs, uint32_t n_jobs, si_fb_batch** out);
/* Re-plans an existing batch for other jobs, keeping its device tables,
 * staging, streams and events where they have room (it allocates only when a
 * class outgrows every table the batch holds), so a serving loop can cycle a
 * few batches without allocating per request.  Variants go back to the
 * defaults, untuned.  The batch's previous runs must have completed, or be
 * ordered before its next run on that run's stream.  No counterpart in the
 * reference, which has no batched form. */
int si_fb_batch_reset(si_fb_batch* batch, const si_fb_job* jobs, uint32_t n_jobs);
size_t si_fb_batch_workspace_bytes(const si_fb_batch* batch);
/* use_graph: capture the launches into a hipGraph on first run for this
 * (workspace, stream) and replay it afterwards.  A class without a variant
 * set or tuned is tuned first (si_fb_batch_tune with the defaults), before
 * any capture. */
int si_fb_batch_run(si_fb_batch* batch, void* workspace, size_t workspace_bytes, int use_graph,
                    void* stream);
int si_fb_batch_destroy(si_fb_batch* batch);

/* Batch classes (one per padded-state count and numeric space) and their
 * variants.  The classes of a batch run side by side (a stream each, forked
 * from and joined back into the caller's stream).  A batch tunes each class
 * with its own persistent kernels on its own jobs (cut to max_markers
 * markers, 0: whole jobs), outputs disabled: every candidate is warmed on
 * 16-marker cuts, then timed twice per pass (the faster run counts), about
 * ten runs of the class in all.  Beside other classes a class keeps the
 * default chain lanes (Lp / 4), and other block lanes only when they are at
 * least 10 % faster alone.  Results are kept per process, keyed by (device,
 * Lp, numeric space, per-marker transitions, size bucket of the class's
 * chain count, alone or beside others), so later batches of the same kind
 * reuse them without timing.  Synchronizes the stream. */
uint32_t si_fb_batch_n_classes(const si_fb_batch* batch);
int si_fb_batch_class_info(const si_fb_batch* batch, uint32_t k, uint32_t* padded_states,
                           uint32_t* flags, si_fb_variant* variant, int* tuned);
int si_fb_batch_set_variant(si_fb_batch* batch, uint32_t k, const si_fb_variant* variant);
int si_fb_batch_tune(si_fb_batch* batch, void* workspace, size_t workspace_bytes,
                     uint32_t max_markers, void* stream);
/* Forget every tuning kept by this process (later batches time again). */
void si_fb_tune_cache_clear(void);

/* Philox4x32-10 uniform the sampler uses for (sample s, marker m): word
 * (m & 3) of Philox4x32-10(counter = {m >> 2, s, rng_stream, 0} with s the
 * global sample index (sample_base + job-local index),
 * key = {seed, seed >> 32}), as (x >> 8) * 2^-24.  Host function. */
float si_sample_uniform(uint64_t seed, uint32_t s, uint32_t m, uint32_t rng_stream);

#ifdef __cplusplus
}
#endif

#endif /* SWIFTIMPUTE_HIP_H */
