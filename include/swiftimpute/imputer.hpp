// SwiftImpute C++ API on MI355X: drop-in for src/api/imputer.hpp of the
// reference (ImputationConfig, ReferencePanel, TargetData, ImputationResult,
// Imputer, free impute()), implemented over the C ABI of
// include/swiftimpute_hip.h.  Same names, fields, defaults and error
// behaviour; additions are marked "added".
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "types.hpp"

namespace swiftimpute {

// Which transition interval each forward-backward step uses.
enum class TransitionMode : uint32_t {
    FirstInterval = 0,  // the reference's kernels: interval 0 for every step
                        // (src/kernels/forward_backward.cu:108)
    PerMarker = 1       // interval m-1 for the step into marker m (added)
};

// Numeric form of the forward-backward kernel.
enum class Precision : uint32_t {
    Auto = 0,       // linear FP32 unless a switch probability is below its range
    LinearF32 = 1,  // per-marker-rescaled linear FP32, FP64 backward scale
    LogF32 = 2      // log-space FP32 (three-term log-sum-exp per state)
};

// Which L-best state selection builds the state table.
enum class SelectionMode : uint32_t {
    CpuSort = 0,   // the reference's CPU path, std::sort over (divergence, haplotype)
                   // (src/pbwt/pbwt_index.cpp:20-44)
    GpuCompat = 1  // the reference's GPU kernels, launch_select_states
                   // (src/pbwt/pbwt_selector.cu:400-463), which choose differently
};

enum class SamplingMode : uint32_t {
    Philox = 0,  // inverse-CDF draw with Philox4x32-10 (seed, sample, marker)
    Argmax = 1   // most probable state (src/kernels/sampling.cu:138-182)
};

// src/api/imputer.hpp:21-56
struct ImputationConfig {
    HMMParameters hmm_params;

    int device_id;               // GPU device (-1: device 0)
    uint32_t batch_size;         // progress is reported every batch_size samples; launches take
                                 // as many multiples of it as fit a 16 GiB device budget
    uint32_t num_threads;        // host threads for input conversion
    bool use_pinned_memory;      // pinned host staging for transfers

    bool output_dosages;         // fill posterior dosages (DS)
    bool output_probabilities;   // fill allele probabilities (AP -> GP)
    bool output_info_score;      // compute INFO scores
    bool deterministic;          // every draw is a function of (seed, sample, marker)

    uint32_t window_size;        // M > window_size > 0: the reference's windowed model
                                 // (imputer.cpp:965-971, 1044-1174); 0: the exact one
    uint32_t window_overlap;
    uint32_t checkpoint_interval;

    // added
    TransitionMode transition_mode;
    SamplingMode sampling;
    Precision precision;
    SelectionMode selection;
    uint64_t seed;               // Philox key (default 12345, sampling.hpp:74)
    bool keep_pbwt_arrays;       // keep prefix/divergence resident after build_index()

    ImputationConfig()
        : device_id(-1),
          batch_size(100),
          num_threads(4),
          use_pinned_memory(true),
          output_dosages(true),
          output_probabilities(true),
          output_info_score(true),
          deterministic(false),
          window_size(10000),
          window_overlap(100),
          checkpoint_interval(0),
          transition_mode(TransitionMode::FirstInterval),
          sampling(SamplingMode::Philox),
          precision(Precision::Auto),
          selection(SelectionMode::CpuSort),
          seed(12345ull),
          keep_pbwt_arrays(true) {}
};

// src/api/imputer.hpp:59-104
class ReferencePanel {
  public:
    ReferencePanel() = default;

    // Plain-text VCF with GT fields (phased or unphased; '.' = missing).
    static std::unique_ptr<ReferencePanel> load_vcf(const std::string& filename,
                                                    const std::string& region = "");
    static std::unique_ptr<ReferencePanel> load_binary(const std::string& filename);
    void save_binary(const std::string& filename) const;

    // added: construct from memory.  alleles is [M][H] (values 0, 1, 255).
    static std::unique_ptr<ReferencePanel> from_alleles(std::vector<Marker> markers,
                                                        std::vector<Sample> samples,
                                                        haplotype_t num_haplotypes,
                                                        const allele_t* alleles);
    // added: construct from bit-packed columns [M][ceil(H/64)] (no missing).
    static std::unique_ptr<ReferencePanel> from_bitpacked(std::vector<Marker> markers,
                                                          haplotype_t num_haplotypes,
                                                          const uint64_t* cols);

    marker_t num_markers() const { return static_cast<marker_t>(markers_.size()); }
    haplotype_t num_haplotypes() const { return num_haplotypes_; }
    sample_t num_samples() const { return num_haplotypes_ / 2; }

    const std::vector<Marker>& markers() const { return markers_; }
    const std::vector<Sample>& samples() const { return samples_; }

    // [M][H] alleles (materialized on first use when built from bit-packed data)
    const allele_t* haplotypes() const;
    allele_t get_allele(marker_t m, haplotype_t h) const;

    size_t memory_usage() const;

    // added: the device input layout
    const std::vector<uint64_t>& packed_alleles() const { return cols_; }  // allele != 0
    const std::vector<uint64_t>& packed_missing() const { return miss_; }  // empty if none
    uint32_t words_per_marker() const { return (num_haplotypes_ + 63) / 64; }

  private:
    std::vector<Marker> markers_;
    std::vector<Sample> samples_;
    haplotype_t num_haplotypes_ = 0;
    std::vector<uint64_t> cols_;
    std::vector<uint64_t> miss_;
    mutable std::vector<allele_t> alleles_;
    void pack_from(const allele_t* alleles);
};

// src/api/imputer.hpp:107-144
class TargetData {
  public:
    TargetData() = default;
    static std::unique_ptr<TargetData> load_vcf(const std::string& filename,
                                                const std::string& region = "");
    // added: GL [B][M] in memory
    static std::unique_ptr<TargetData> from_likelihoods(std::vector<Marker> markers,
                                                        std::vector<Sample> samples,
                                                        const GenotypeLikelihoods* gl);
    // added: palette-coded likelihoods; codes [B][M] index palette
    static std::unique_ptr<TargetData> from_codes(std::vector<Marker> markers, uint32_t num_samples,
                                                  std::vector<GenotypeLikelihoods> palette,
                                                  const uint8_t* codes);

    marker_t num_markers() const { return static_cast<marker_t>(markers_.size()); }
    sample_t num_samples() const { return static_cast<sample_t>(samples_.size()); }
    const std::vector<Marker>& markers() const { return markers_; }
    const std::vector<Sample>& samples() const { return samples_; }

    // [B][M] likelihoods (materialized on first use when built from codes)
    const GenotypeLikelihoods* genotype_likelihoods() const;
    GenotypeLikelihoods get_likelihood(sample_t s, marker_t m) const;
    size_t memory_usage() const;

    // added: the device input layout (distinct GL triples + one code per
    // (sample, marker), rows padded to a multiple of 16)
    const std::vector<GenotypeLikelihoods>& palette() const { return palette_; }
    const std::vector<uint8_t>& codes() const { return codes_; }
    uint32_t code_pitch() const { return pitch_; }

  private:
    std::vector<Marker> markers_;
    std::vector<Sample> samples_;
    std::vector<GenotypeLikelihoods> palette_;
    std::vector<uint8_t> codes_;
    uint32_t pitch_ = 0;
    mutable std::vector<GenotypeLikelihoods> gl_;
    void encode(const GenotypeLikelihoods* gl);
};

// src/api/imputer.hpp:147-194
class ImputationResult {
  public:
    ImputationResult(uint32_t num_samples, uint32_t num_markers);

    void write_vcf(const std::string& filename, const TargetData& targets,
                   const ReferencePanel& reference, const ImputationConfig& config) const;

    uint32_t num_samples() const { return num_samples_; }
    uint32_t num_markers() const { return num_markers_; }

    const allele_t* haplotype0(sample_t s) const;
    const allele_t* haplotype1(sample_t s) const;

    // sampled-allele dosage h0 + h1 (src/api/imputer.cpp:418-425)
    prob_t get_dosage(sample_t s, marker_t m) const;
    // from the posterior allele probabilities (src/api/imputer.cpp:427-445)
    void get_probabilities(sample_t s, marker_t m, prob_t probs[3]) const;
    prob_t get_info_score(marker_t m) const;

    void set_haplotype(sample_t s, uint32_t phase, const allele_t* haplotype);
    void set_posterior(sample_t s, marker_t m, const prob_t* posterior);
    void compute_info_scores();
    size_t memory_usage() const;

    // added: posterior dosage AP0 + AP1 and raw access.  has_posteriors()
    // says whether the allele probabilities were filled (set_posterior, or an
    // Imputer run with output_dosages or output_probabilities); INFO and the
    // VCF's DS/GP/AP fields use them when they are.
    prob_t get_posterior_dosage(sample_t s, marker_t m) const;
    bool has_posteriors() const { return has_posteriors_; }
    void set_has_posteriors(bool v) { has_posteriors_ = v; }
    allele_t* mutable_haplotypes() { return haplotypes_.get(); }
    prob_t* mutable_posteriors() { return posteriors_.get(); }
    double log_likelihood(sample_t s) const { return s < loglik_.size() ? loglik_[s] : 0.0; }
    std::vector<double>& mutable_log_likelihoods() { return loglik_; }

  private:
    uint32_t num_samples_;
    uint32_t num_markers_;
    std::unique_ptr<allele_t[]> haplotypes_;  // [B][2][M]
    std::unique_ptr<prob_t[]> posteriors_;    // [B][M][2] (AP0, AP1)
    std::unique_ptr<prob_t[]> info_scores_;   // [M]
    std::vector<double> loglik_;
    bool has_posteriors_ = false;
};

// src/api/imputer.hpp:197-287
class Imputer {
  public:
    Imputer(const ReferencePanel& reference, const ImputationConfig& config = ImputationConfig());
    ~Imputer();
    Imputer(const Imputer&) = delete;
    Imputer& operator=(const Imputer&) = delete;

    void build_index();
    bool has_index() const { return index_built_; }

    std::unique_ptr<ImputationResult> impute(const TargetData& targets);
    std::unique_ptr<ImputationResult> impute_with_progress(
        const TargetData& targets, std::function<void(uint32_t, uint32_t)> progress_callback);

    const ImputationConfig& config() const { return config_; }
    size_t device_memory_usage() const;
    static size_t estimate_memory_requirement(const ReferencePanel& reference,
                                              uint32_t num_target_samples,
                                              const ImputationConfig& config);

    // added: PBWT arrays and state table, copied to the host
    std::vector<uint32_t> pbwt_prefix() const;      // [M][H]
    std::vector<uint32_t> pbwt_divergence() const;  // [M][H]
    std::vector<uint32_t> selected_states() const;  // [M][L]

  private:
    struct Device;
    const ReferencePanel& reference_;
    ImputationConfig config_;
    std::unique_ptr<Device> dev_;
    bool index_built_ = false;
    void validate_targets(const TargetData& targets) const;
    void impute_batch(const TargetData& targets, uint32_t start, uint32_t end,
                      ImputationResult& result);
};

// src/api/imputer.hpp:290-295
std::unique_ptr<ImputationResult> impute(const std::string& reference_vcf,
                                         const std::string& target_vcf,
                                         const std::string& output_vcf,
                                         const ImputationConfig& config = ImputationConfig());

}  // namespace swiftimpute
