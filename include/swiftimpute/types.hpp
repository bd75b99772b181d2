// Core types of the SwiftImpute C++ API (drop-in for src/core/types.hpp of
// the reference: same aliases, constants, structs and exception hierarchy,
// src/core/types.hpp:12-90,222-237).  Device errors are DeviceError (HIP);
// CUDAError remains as a name for it so existing catch clauses compile.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace swiftimpute {

using marker_t = uint32_t;     // marker index
using sample_t = uint32_t;     // sample index
using haplotype_t = uint32_t;  // haplotype index (0 .. 2N-1)
using allele_t = uint8_t;      // 0, 1 or ALLELE_MISSING
using state_t = uint16_t;      // HMM state index
using prob_t = float;

constexpr allele_t ALLELE_MISSING = 255;
constexpr marker_t INVALID_MARKER = 0xFFFFFFFF;
constexpr sample_t INVALID_SAMPLE = 0xFFFFFFFF;

struct Marker {
    std::string chrom;
    uint64_t pos;
    std::string id;
    std::string ref;
    std::string alt;
    double cM;  // genetic position; 0 = derive from pos at 1 cM/Mb
    Marker() : pos(0), cM(0.0) {}
};

struct Sample {
    std::string id;
    uint32_t index;
    Sample() : index(0) {}
    explicit Sample(std::string id_) : id(std::move(id_)), index(0) {}
};

// Genotype likelihoods, log10 scale.
struct GenotypeLikelihoods {
    prob_t ll_00;
    prob_t ll_01;
    prob_t ll_11;
    GenotypeLikelihoods() : ll_00(0.0f), ll_01(0.0f), ll_11(0.0f) {}
    GenotypeLikelihoods(prob_t a, prob_t b, prob_t c) : ll_00(a), ll_01(b), ll_11(c) {}
};

struct PhasedHaplotype {
    allele_t* hap0;
    allele_t* hap1;
    prob_t* probabilities;
    PhasedHaplotype() : hap0(nullptr), hap1(nullptr), probabilities(nullptr) {}
};

struct QualityMetrics {
    prob_t info_score;
    prob_t r2_score;
    prob_t concordance;
    uint32_t num_called;
    uint32_t num_missing;
    QualityMetrics() : info_score(0.0f), r2_score(0.0f), concordance(0.0f), num_called(0), num_missing(0) {}
};

// Li-Stephens model parameters (defaults as src/core/types.hpp:84-89).
struct HMMParameters {
    prob_t theta;             // mutation rate (unused by the model, as in the reference)
    prob_t rho_rate;          // recombination multiplier
    uint32_t ne;              // effective population size
    uint32_t num_states;      // L
    bool use_pbwt_selection;  // PBWT state selection (else pairs (2k, 2k+1))
    HMMParameters() : theta(0.001f), rho_rate(4.0f), ne(10000), num_states(8), use_pbwt_selection(true) {}
};

class ImputationError : public std::runtime_error {
  public:
    explicit ImputationError(const std::string& msg) : std::runtime_error(msg) {}
};

// A failing HIP runtime call or kernel launch (the reference's CUDAError).
class DeviceError : public ImputationError {
  public:
    DeviceError(const std::string& msg, int error_code)
        : ImputationError(msg + " (device error: " + std::to_string(error_code) + ")"), error_code_(error_code) {}
    int error_code() const { return error_code_; }

  private:
    int error_code_;
};
using CUDAError = DeviceError;

// A shape or mode this build does not handle (SI_ERR_UNSUPPORTED from the C
// ABI).  No counterpart in the reference; it derives from ImputationError so
// that callers catching the reference's hierarchy still see it.
class UnsupportedError : public ImputationError {
  public:
    explicit UnsupportedError(const std::string& msg) : ImputationError(msg) {}
};

// Minimal leveled logger (the reference's Logger, src/core/types.hpp:253-283).
class Logger {
  public:
    enum Level { DEBUG = 0, INFO = 1, WARNING = 2, ERROR = 3 };
    static Logger& instance();
    void set_level(Level level) { level_ = level; }
    Level get_level() const { return level_; }
    void debug(const std::string& msg) { log(DEBUG, "DEBUG", msg); }
    void info(const std::string& msg) { log(INFO, "INFO", msg); }
    void warning(const std::string& msg) { log(WARNING, "WARNING", msg); }
    void error(const std::string& msg) { log(ERROR, "ERROR", msg); }

  private:
    Logger() : level_(WARNING) {}
    Level level_;
    void log(Level level, const char* prefix, const std::string& msg);
};

}  // namespace swiftimpute
