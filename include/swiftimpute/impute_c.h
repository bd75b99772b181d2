/*
 * impute_c.h -- a C entry point of libswiftimpute.so for FFI callers: the
 * whole imputation of one VCF cohort against one VCF panel, with the stage
 * timings the reference's executable reports.
 *
 * si_impute_vcf restates the reference's free function
 * impute(reference_vcf, target_vcf, output_vcf, config)
 * (src/api/imputer.cpp:1241-1270) over the C++ API of
 * include/swiftimpute/imputer.hpp, timed per stage as src/main.cpp:96-216
 * does (reference load, target load, PBWT index, imputation, output).
 */
#ifndef SWIFTIMPUTE_IMPUTE_C_H
#define SWIFTIMPUTE_IMPUTE_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct si_impute_opts {
    uint32_t num_states;   /* L (HMMParameters::num_states, default 8) */
    uint32_t ne;           /* effective population size (10000) */
    uint32_t batch_size;   /* ImputationConfig::batch_size (100) */
    uint32_t window_size;  /* ImputationConfig::window_size (10000; 0: exact model) */
    uint32_t precision;    /* 0 Auto, 1 LinearF32, 2 LogF32 (Precision) */
    uint32_t per_marker;   /* 0: first interval's transitions (the reference), 1: per marker */
    uint64_t seed;         /* Philox key (12345) */
} si_impute_opts;

typedef struct si_impute_timing {
    double reference_load_s, target_load_s, index_s, impute_s, write_s, total_s;
    uint64_t samples, markers, untyped_genotypes;
} si_impute_timing;

void si_impute_defaults(si_impute_opts* opts);

/* Loads both VCFs, builds the index, imputes and writes output_vcf.  Returns
 * 0 or a negative status (-1 argument/input error, -2 device error, -3
 * unsupported shape) with the message in si_impute_last_error(). */
int si_impute_vcf(const char* reference_vcf, const char* target_vcf, const char* output_vcf,
                  const si_impute_opts* opts, si_impute_timing* timing);
const char* si_impute_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SWIFTIMPUTE_IMPUTE_C_H */
