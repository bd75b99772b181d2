/*
 * synth.h -- seeded synthetic reference panels and target cohorts (C ABI of
 * libswiftimpute.so).  The benchmark workloads are synthetic: panels are
 * mosaics of founder haplotypes, targets are mosaics of panel haplotypes with
 * genotyping error, typed at a regular subset of markers.  All randomness is
 * Philox4x32-10 keyed by the seed and indexed by (purpose, item), so output
 * is identical for any thread count.
 */
#ifndef SWIFTIMPUTE_SYNTH_H
#define SWIFTIMPUTE_SYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct si_synth_panel_cfg {
    uint32_t n_haplotypes;    /* H */
    uint32_t n_markers;       /* M */
    uint32_t n_founders;      /* F (default 40), at most 64 */
    uint32_t spacing_mode;    /* 0: constant spacing_bp; 1: log-normal spacing + hotspots */
    uint64_t seed;
    double mutation;          /* per-allele flip probability (1e-3) */
    double spacing_bp;        /* mode 0 spacing (50 kb) */
    double lognorm_median_bp; /* mode 1 median spacing (20 kb) */
    double lognorm_sigma;     /* mode 1 sigma (1.5) */
    double hotspot_frac;      /* mode 1 fraction of intervals that are hotspots (0.01) */
    double hotspot_mult;      /* mode 1 hotspot rate multiplier (50) */
} si_synth_panel_cfg;

/* Fills a cfg with the defaults above for H haplotypes x M markers. */
void si_synth_panel_defaults(si_synth_panel_cfg* cfg, uint32_t H, uint32_t M, uint64_t seed);

/* cols [M][ceil(H/64)] bit-packed alleles; pos [M] base pairs; cM [M]
 * genetic positions in centiMorgans (1 cM/Mb background; hotspots in mode 1);
 * morgans [M] cumulative genetic map used for the copying process.  Any
 * output except cols may be NULL.  Returns 0 or a negative status. */
int si_synth_panel(const si_synth_panel_cfg* cfg, uint64_t* cols, uint64_t* pos, double* cM,
                   double* morgans, int n_threads);

typedef struct si_synth_target_cfg {
    uint32_t n_samples;       /* B */
    uint32_t typed_every;     /* markers with m % typed_every == typed_offset are typed */
    uint32_t typed_offset;
    uint32_t code_pitch;      /* row pitch of codes (>= M) */
    uint64_t seed;
    double het_lo, het_hi;    /* per-sample heterozygosity target, uniform (0.12, 0.40) */
    double geno_error;        /* typed-genotype error rate (0.005) */
} si_synth_target_cfg;

void si_synth_target_defaults(si_synth_target_cfg* cfg, uint32_t B, uint32_t M, uint32_t typed_every,
                              uint64_t seed);

/* Targets against a panel made by si_synth_panel (cols, H, M, morgans).
 * codes [B][code_pitch]: 0 = 0/0, 1 = 0/1, 2 = 1/1 (hard calls at typed
 * markers), 3 = missing (untyped).  truth [B][2][M] (may be NULL) holds the
 * true haplotypes.  Returns 0 or a negative status. */
int si_synth_targets(const si_synth_target_cfg* cfg, const uint64_t* cols, uint32_t H, uint32_t M,
                     const double* morgans, uint8_t* codes, uint8_t* truth, int n_threads);

/* The genotype-likelihood palette of hard calls, in the order of the codes
 * above: log10 (ll_00, ll_01, ll_11) = (0,-999,-999), (-999,0,-999),
 * (-999,-999,0), (0,0,0) -- the conversion of TargetData::load_vcf
 * (src/api/imputer.cpp:244-293). */
void si_hardcall_palette(float palette[12]);

#ifdef __cplusplus
}
#endif

#endif /* SWIFTIMPUTE_SYNTH_H */
