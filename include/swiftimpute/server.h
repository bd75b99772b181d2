/*
 * server.h -- the serving form (C ABI of libswiftimpute.so).
 *
 * A server keeps one reference panel resident in HBM (bit-packed alleles, the
 * PBWT-selected state table and the per-marker transition coefficients) and
 * imputes requests -- a target cohort over a contiguous marker region of the
 * panel -- as they arrive.  A scheduler thread admits queued requests in
 * arrival order while their device footprint (codes, outputs, checkpoints)
 * fits the HBM budget, forms a batch of every admitted request, and launches
 * it (si_fb_batch_*) on one of several streams, so requests that arrive while
 * others run are admitted as memory frees up and their work overlaps on the
 * device.  Each request's result depends only on its own inputs and seed:
 * alone or in any mix it is bit-for-bit the same.
 *
 * Region semantics: a request over [m0, m1) uses the resident full-panel
 * PBWT order at each marker of the region (the reference's --region path
 * re-sweeps the PBWT from the region start instead).
 *
 * No counterpart in the reference (its Imputer processes one cohort file at
 * a time, src/api/imputer.cpp:747-785).
 */
#ifndef SWIFTIMPUTE_SERVER_H
#define SWIFTIMPUTE_SERVER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct si_server si_server;

typedef struct si_server_cfg {
    uint32_t n_states;          /* L */
    uint32_t ne;                /* effective population size */
    float rho_rate;             /* recombination multiplier (4.0) */
    uint32_t per_marker;        /* 0: each request uses its region's first interval */
    uint32_t max_streams;       /* batches in flight at once (>= 1) */
    uint32_t max_batch;         /* requests per batch (>= 1) */
    uint64_t hbm_budget_bytes;  /* device memory for admitted requests */
    int device_id;
    uint32_t max_queue_delay_us; /* a partial batch waits up to this long for company (0) */
    uint32_t selection;         /* 0: CPU-path selection (pbwt_index.cpp:20-44);
                                   1: the reference's GPU kernels (si_select_perm_gpu_compat) */
    uint32_t region_resweep;    /* 0: a request over [m0, m1) uses the resident panel-wide
                                   PBWT rows; 1: its state table comes from a PBWT sweep
                                   of the region alone, started from the identity order
                                   at m0, as the reference's --region run does
                                   (main.cpp:103, imputer.cpp:40-44); costs a sweep and
                                   M*H*4 bytes per request */
    uint32_t plain;             /* 0: a slot's requests run as one batched launch per pass;
                                   1: each request alone through si_fb_run (one job per
                                   launch), the same staging and copies otherwise */
    uint64_t staging_bytes;     /* page-locked host pool for queued requests' codes (4 GiB):
                                   submit copies the codes into it and the upload reads
                                   them from it; a request that finds it full keeps its
                                   codes in pageable memory, uploaded through a staging
                                   copy (0: no pool) */
    uint32_t window_size;       /* the reference's windowed model for requests longer than
                                   this many markers (imputer.cpp:965-971; its default
                                   configuration uses 10,000); 0: the exact model for every
                                   request.  At least 32 when set. */
} si_server_cfg;

void si_server_defaults(si_server_cfg* cfg);

/* cols [M][ceil(H/64)] host bit-packed panel; pos [M] bp; cM [M] or NULL.
 * Builds the PBWT index and state table on the device, then starts the
 * scheduler thread.  Returns 0 or a negative status (si_server_last_error). */
int si_server_create(const si_server_cfg* cfg, const uint64_t* cols, uint32_t H, uint32_t M,
                     const uint64_t* pos, const double* cM, si_server** out);

/* Queue a request: cohort of B samples over markers [m0, m1) with hard-call
 * codes [B][m1-m0] (0 = 0/0, 1 = 0/1, 2 = 1/1, 3 = missing), Philox seed.
 * codes is copied; the call returns at once with the request id. */
int si_server_submit(si_server* s, uint32_t m0, uint32_t m1, uint32_t B, const uint8_t* codes,
                     uint64_t seed, uint64_t* request_id);

/* The same, with the results written straight into caller buffers hap
 * [B][2][m1-m0] and ds [B][m1-m0] (either may be NULL), which must stay valid
 * until si_server_wait returns for this request; saves one host copy.  A
 * buffer in page-locked memory (hipHostMalloc) receives the device copy
 * itself, without passing through the server's staging. */
int si_server_submit_into(si_server* s, uint32_t m0, uint32_t m1, uint32_t B,
                          const uint8_t* codes, uint64_t seed, uint8_t* hap, float* ds,
                          uint64_t* request_id);

/* Wait for a request and copy its results: hap [B][2][M'] and ds [B][M']
 * (either may be NULL; ignored for si_server_submit_into requests, whose
 * results are already in place); latency_s receives arrival-to-result
 * seconds. */
int si_server_wait(si_server* s, uint64_t request_id, uint8_t* hap, float* ds, double* latency_s);

/* Counters since creation: completed requests, batches, imputed genotypes
 * (untyped (sample, marker) pairs), largest batch, peak admitted bytes,
 * admissions deferred for memory, and requests whose results the device
 * copied straight into page-locked caller buffers (si_server_submit_into).
 * Where the time goes, summed over retired batches: bytes copied to and
 * from the device; the scheduler's host time staging and enqueueing them;
 * each batch's stream spans for its code upload, its forward-backward
 * launches and its result copies (batches on different streams overlap, so
 * these can add up to more than the wall time); the copier threads' host
 * time moving results out of staging. */
typedef struct si_server_stats {
    uint64_t requests, batches, genotypes, untyped_genotypes, max_batch_requests;
    uint64_t peak_bytes, deferred_admissions, direct_copies;
    uint64_t h2d_bytes, d2h_bytes;
    double launch_host_s, h2d_s, kernel_s, d2h_s, copy_host_s;
} si_server_stats;
int si_server_get_stats(si_server* s, si_server_stats* out);

int si_server_destroy(si_server* s);
const char* si_server_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SWIFTIMPUTE_SERVER_H */
