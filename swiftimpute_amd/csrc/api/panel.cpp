// ReferencePanel, TargetData, ImputationResult and plain-text VCF I/O of the
// C++ API (include/swiftimpute/imputer.hpp).
//
// Reference behaviour kept: panel haplotype 2s+p holds allele p of sample s
// (src/api/imputer.cpp:85-112); target genotypes become hard-call log10
// likelihoods 0 / -999, missing or non-biallelic calls (0, 0, 0)
// (imputer.cpp:244-293); write_vcf emits phased GT lines
// (imputer.cpp:354-406 -> vcf_writer.hpp:212-250).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>

#include "../../../include/swiftimpute/imputer.hpp"

namespace swiftimpute {

// ------------------------------------------------------------------ Logger
Logger& Logger::instance() {
    static Logger l;
    return l;
}

void Logger::log(Level level, const char* prefix, const std::string& msg) {
    static std::mutex mu;
    if (level < level_) return;
    std::lock_guard<std::mutex> g(mu);
    std::cout << "[" << prefix << "] " << msg << std::endl;
}

// --------------------------------------------------------------- VCF input
namespace {

struct VcfRecord {
    Marker marker;
    std::vector<std::vector<allele_t>> gts;  // per sample
};

struct Region {
    bool any = true;
    std::string chrom;
    uint64_t lo = 0, hi = UINT64_MAX;
};

Region parse_region(const std::string& r) {
    Region g;
    if (r.empty()) return g;
    g.any = false;
    const auto c = r.find(':');
    g.chrom = r.substr(0, c);
    if (c != std::string::npos) {
        const std::string span = r.substr(c + 1);
        const auto d = span.find('-');
        g.lo = std::stoull(span.substr(0, d));
        if (d != std::string::npos && d + 1 < span.size()) g.hi = std::stoull(span.substr(d + 1));
    }
    return g;
}

std::vector<std::string> split(const std::string& s, char sep) {
    std::vector<std::string> out;
    size_t a = 0;
    for (;;) {
        const size_t b = s.find(sep, a);
        out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
        if (b == std::string::npos) break;
        a = b + 1;
    }
    return out;
}

std::vector<allele_t> parse_gt(const std::string& field, int gt_index) {
    std::vector<allele_t> out;
    const auto parts = split(field, ':');
    if (gt_index < 0 || size_t(gt_index) >= parts.size()) return out;
    const std::string& gt = parts[size_t(gt_index)];
    size_t a = 0;
    for (size_t i = 0; i <= gt.size(); ++i) {
        if (i == gt.size() || gt[i] == '|' || gt[i] == '/') {
            const std::string tok = gt.substr(a, i - a);
            if (tok.empty() || tok == ".") out.push_back(ALLELE_MISSING);
            else out.push_back(static_cast<allele_t>(std::min(254, std::stoi(tok))));
            a = i + 1;
        }
    }
    return out;
}

void read_vcf(const std::string& filename, const std::string& region, std::vector<std::string>& names,
              std::vector<VcfRecord>& recs) {
    if (filename.size() > 3 &&
        (filename.compare(filename.size() - 3, 3, ".gz") == 0 ||
         filename.compare(filename.size() - 4, 4, ".bcf") == 0))
        throw ImputationError("compressed VCF/BCF input needs htslib, which this build does not use: " +
                              filename);
    std::ifstream in(filename);
    if (!in) throw ImputationError("cannot open VCF file: " + filename);
    const Region rg = parse_region(region);
    std::string line;
    bool header = false;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        if (line.rfind("##", 0) == 0) continue;
        if (line[0] == '#') {
            const auto cols = split(line, '\t');
            for (size_t i = 9; i < cols.size(); ++i) names.push_back(cols[i]);
            header = true;
            continue;
        }
        if (!header) throw ImputationError("VCF data line before the #CHROM header: " + filename);
        const auto cols = split(line, '\t');
        if (cols.size() < 8) throw ImputationError("malformed VCF line in " + filename);
        VcfRecord r;
        r.marker.chrom = cols[0];
        r.marker.pos = std::stoull(cols[1]);
        if (!rg.any && (r.marker.chrom != rg.chrom || r.marker.pos < rg.lo || r.marker.pos > rg.hi))
            continue;
        r.marker.id = cols[2] == "." ? "" : cols[2];
        r.marker.ref = cols[3];
        r.marker.alt = split(cols[4], ',')[0];
        int gt_index = -1;
        if (cols.size() > 8) {
            const auto fmt = split(cols[8], ':');
            for (size_t i = 0; i < fmt.size(); ++i)
                if (fmt[i] == "GT") gt_index = int(i);
        }
        for (size_t i = 9; i < cols.size(); ++i) r.gts.push_back(parse_gt(cols[i], gt_index));
        recs.push_back(std::move(r));
    }
}

std::vector<Sample> make_samples(const std::vector<std::string>& names) {
    std::vector<Sample> out;
    for (size_t i = 0; i < names.size(); ++i) {
        Sample s(names[i]);
        s.index = uint32_t(i);
        out.push_back(s);
    }
    return out;
}

}  // namespace

// ------------------------------------------------------------ ReferencePanel
void ReferencePanel::pack_from(const allele_t* a) {
    const size_t M = markers_.size(), H = num_haplotypes_, W = (H + 63) / 64;
    cols_.assign(M * W, 0);
    bool any_missing = false;
    for (size_t m = 0; m < M; ++m)
        for (size_t h = 0; h < H; ++h) {
            const allele_t v = a[m * H + h];
            if (v != 0 && v != 1 && v != ALLELE_MISSING)
                throw ImputationError("allele value " + std::to_string(int(v)) + " at marker " +
                                      std::to_string(m) + ": only biallelic panels (0/1/missing) are supported");
            if (v != 0) cols_[m * W + h / 64] |= 1ull << (h % 64);
            any_missing |= (v == ALLELE_MISSING);
        }
    miss_.clear();
    if (any_missing) {
        miss_.assign(M * W, 0);
        for (size_t m = 0; m < M; ++m)
            for (size_t h = 0; h < H; ++h)
                if (a[m * H + h] == ALLELE_MISSING) miss_[m * W + h / 64] |= 1ull << (h % 64);
    }
}

std::unique_ptr<ReferencePanel> ReferencePanel::load_vcf(const std::string& filename,
                                                         const std::string& region) {
    std::vector<std::string> names;
    std::vector<VcfRecord> recs;
    read_vcf(filename, region, names, recs);
    if (recs.empty()) throw ImputationError("No variants found in reference VCF");
    const size_t M = recs.size(), S = names.size(), H = 2 * S;
    std::vector<allele_t> a(M * H, ALLELE_MISSING);
    std::vector<Marker> markers;
    for (size_t m = 0; m < M; ++m) {
        markers.push_back(recs[m].marker);
        for (size_t s = 0; s < S; ++s) {
            if (s >= recs[m].gts.size()) continue;
            const auto& g = recs[m].gts[s];
            if (g.size() >= 2) {
                a[m * H + 2 * s] = g[0];
                a[m * H + 2 * s + 1] = g[1];
            } else if (g.size() == 1) {
                a[m * H + 2 * s] = a[m * H + 2 * s + 1] = g[0];
            }
        }
    }
    return from_alleles(std::move(markers), make_samples(names), haplotype_t(H), a.data());
}

std::unique_ptr<ReferencePanel> ReferencePanel::load_binary(const std::string&) {
    throw ImputationError("Binary reference format not yet implemented");
}

void ReferencePanel::save_binary(const std::string&) const {
    throw ImputationError("Binary reference format not yet implemented");
}

std::unique_ptr<ReferencePanel> ReferencePanel::from_alleles(std::vector<Marker> markers,
                                                             std::vector<Sample> samples,
                                                             haplotype_t H, const allele_t* alleles) {
    if (H == 0) throw ImputationError("reference panel has no haplotypes");
    auto p = std::make_unique<ReferencePanel>();
    p->markers_ = std::move(markers);
    p->samples_ = std::move(samples);
    p->num_haplotypes_ = H;
    p->pack_from(alleles);
    return p;
}

std::unique_ptr<ReferencePanel> ReferencePanel::from_bitpacked(std::vector<Marker> markers,
                                                               haplotype_t H, const uint64_t* cols) {
    if (H == 0) throw ImputationError("reference panel has no haplotypes");
    auto p = std::make_unique<ReferencePanel>();
    const size_t W = (size_t(H) + 63) / 64;
    p->cols_.assign(cols, cols + markers.size() * W);
    p->markers_ = std::move(markers);
    p->num_haplotypes_ = H;
    for (uint32_t s = 0; s < H / 2; ++s) {
        Sample x("HAP_SAMPLE_" + std::to_string(s));
        x.index = s;
        p->samples_.push_back(x);
    }
    return p;
}

const allele_t* ReferencePanel::haplotypes() const {
    if (alleles_.empty() && !markers_.empty()) {
        const size_t M = markers_.size(), H = num_haplotypes_, W = words_per_marker();
        alleles_.resize(M * H);
        for (size_t m = 0; m < M; ++m)
            for (size_t h = 0; h < H; ++h) {
                const uint64_t bit = 1ull << (h % 64);
                allele_t v = (cols_[m * W + h / 64] & bit) ? 1 : 0;
                if (!miss_.empty() && (miss_[m * W + h / 64] & bit)) v = ALLELE_MISSING;
                alleles_[m * H + h] = v;
            }
    }
    return alleles_.data();
}

allele_t ReferencePanel::get_allele(marker_t m, haplotype_t h) const {
    if (m >= markers_.size() || h >= num_haplotypes_) return ALLELE_MISSING;
    const size_t W = words_per_marker();
    const uint64_t bit = 1ull << (h % 64);
    if (!miss_.empty() && (miss_[size_t(m) * W + h / 64] & bit)) return ALLELE_MISSING;
    return (cols_[size_t(m) * W + h / 64] & bit) ? 1 : 0;
}

size_t ReferencePanel::memory_usage() const {
    size_t total = markers_.capacity() * sizeof(Marker) + samples_.capacity() * sizeof(Sample);
    for (const auto& m : markers_) total += m.chrom.capacity() + m.id.capacity() + m.ref.capacity() + m.alt.capacity();
    for (const auto& s : samples_) total += s.id.capacity();
    total += (cols_.size() + miss_.size()) * sizeof(uint64_t) + alleles_.size();
    return total;
}

// ---------------------------------------------------------------- TargetData
namespace {
const prob_t kLog10Zero = -999.0f;

GenotypeLikelihoods hard_call(const std::vector<allele_t>& g) {
    if (g.size() < 2 || g[0] == ALLELE_MISSING || g[1] == ALLELE_MISSING) return {0.f, 0.f, 0.f};
    if (g[0] == 0 && g[1] == 0) return {0.f, kLog10Zero, kLog10Zero};
    if ((g[0] == 0 && g[1] == 1) || (g[0] == 1 && g[1] == 0)) return {kLog10Zero, 0.f, kLog10Zero};
    if (g[0] == 1 && g[1] == 1) return {kLog10Zero, kLog10Zero, 0.f};
    return {0.f, 0.f, 0.f};
}
}  // namespace

void TargetData::encode(const GenotypeLikelihoods* gl) {
    const size_t B = samples_.size(), M = markers_.size();
    pitch_ = uint32_t((M + 15) / 16 * 16);
    codes_.assign(B * pitch_, 0);
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint8_t> index;
    palette_.clear();
    auto key = [](const GenotypeLikelihoods& g) {
        uint32_t a, b, c;
        std::memcpy(&a, &g.ll_00, 4);
        std::memcpy(&b, &g.ll_01, 4);
        std::memcpy(&c, &g.ll_11, 4);
        return std::make_tuple(a, b, c);
    };
    for (size_t s = 0; s < B; ++s)
        for (size_t m = 0; m < M; ++m) {
            const auto& g = gl[s * M + m];
            const auto k = key(g);
            auto it = index.find(k);
            if (it == index.end()) {
                if (palette_.size() >= 64)
                    throw ImputationError(
                        "more than 64 distinct genotype-likelihood triples: this build keeps the "
                        "likelihood palette in on-chip memory and supports at most 64");
                it = index.emplace(k, uint8_t(palette_.size())).first;
                palette_.push_back(g);
            }
            codes_[s * pitch_ + m] = it->second;
        }
    if (palette_.empty()) palette_.push_back(GenotypeLikelihoods());
    for (size_t s = 0; s < B; ++s)
        for (size_t m = M; m < pitch_; ++m) codes_[s * pitch_ + m] = 0;
}

std::unique_ptr<TargetData> TargetData::load_vcf(const std::string& filename, const std::string& region) {
    std::vector<std::string> names;
    std::vector<VcfRecord> recs;
    read_vcf(filename, region, names, recs);
    if (recs.empty()) throw ImputationError("No variants found in target VCF");
    const size_t M = recs.size(), B = names.size();
    std::vector<GenotypeLikelihoods> gl(B * M);
    std::vector<Marker> markers;
    for (size_t m = 0; m < M; ++m) {
        markers.push_back(recs[m].marker);
        for (size_t s = 0; s < B; ++s)
            gl[s * M + m] = s < recs[m].gts.size() ? hard_call(recs[m].gts[s]) : GenotypeLikelihoods();
    }
    return from_likelihoods(std::move(markers), make_samples(names), gl.data());
}

std::unique_ptr<TargetData> TargetData::from_likelihoods(std::vector<Marker> markers,
                                                         std::vector<Sample> samples,
                                                         const GenotypeLikelihoods* gl) {
    auto t = std::make_unique<TargetData>();
    t->markers_ = std::move(markers);
    t->samples_ = std::move(samples);
    t->encode(gl);
    return t;
}

std::unique_ptr<TargetData> TargetData::from_codes(std::vector<Marker> markers, uint32_t B,
                                                   std::vector<GenotypeLikelihoods> palette,
                                                   const uint8_t* codes) {
    if (palette.empty() || palette.size() > 64)
        throw ImputationError("palette must hold 1..64 genotype-likelihood triples");
    auto t = std::make_unique<TargetData>();
    t->markers_ = std::move(markers);
    for (uint32_t s = 0; s < B; ++s) {
        Sample x("TARGET_" + std::to_string(s));
        x.index = s;
        t->samples_.push_back(x);
    }
    const size_t M = t->markers_.size();
    t->pitch_ = uint32_t((M + 15) / 16 * 16);
    t->codes_.assign(size_t(B) * t->pitch_, 0);
    for (size_t s = 0; s < B; ++s)
        for (size_t m = 0; m < M; ++m) {
            const uint8_t c = codes[s * M + m];
            if (c >= palette.size()) throw ImputationError("genotype code outside the palette");
            t->codes_[s * t->pitch_ + m] = c;
        }
    t->palette_ = std::move(palette);
    return t;
}

const GenotypeLikelihoods* TargetData::genotype_likelihoods() const {
    if (gl_.empty() && !samples_.empty()) {
        const size_t B = samples_.size(), M = markers_.size();
        gl_.resize(B * M);
        for (size_t s = 0; s < B; ++s)
            for (size_t m = 0; m < M; ++m) gl_[s * M + m] = palette_[codes_[s * pitch_ + m]];
    }
    return gl_.data();
}

GenotypeLikelihoods TargetData::get_likelihood(sample_t s, marker_t m) const {
    if (s >= samples_.size() || m >= markers_.size()) return GenotypeLikelihoods();
    return palette_[codes_[size_t(s) * pitch_ + m]];
}

size_t TargetData::memory_usage() const {
    return markers_.capacity() * sizeof(Marker) + samples_.capacity() * sizeof(Sample) + codes_.size() +
           palette_.size() * sizeof(GenotypeLikelihoods) + gl_.size() * sizeof(GenotypeLikelihoods);
}

// ---------------------------------------------------------- ImputationResult
ImputationResult::ImputationResult(uint32_t B, uint32_t M) : num_samples_(B), num_markers_(M) {
    const size_t hap = size_t(B) * 2 * M, post = size_t(B) * M * 2;
    haplotypes_ = std::make_unique<allele_t[]>(hap);
    posteriors_ = std::make_unique<prob_t[]>(post);
    info_scores_ = std::make_unique<prob_t[]>(M);
    std::fill(haplotypes_.get(), haplotypes_.get() + hap, 0);
    std::fill(posteriors_.get(), posteriors_.get() + post, 0.0f);
    std::fill(info_scores_.get(), info_scores_.get() + M, 0.0f);
    loglik_.assign(B, 0.0);
}

// The reference's writer (VCFWriter, src/io/vcf_writer.hpp:97-210; called
// from ImputationResult::write_vcf, src/api/imputer.cpp:354-406) declares
// DS/GP/AP but writes only the phased GT.  Here the FORMAT carries what the
// config asks for and the result holds: GT (the sampled haplotypes, "." for a
// missing panel allele), DS = AP0 + AP1 when output_dosages, GP and AP when
// output_probabilities (4 decimals, the reference's format_dosage /
// format_genotype_probs); INFO carries the score when output_info_score.
// Only the fields written are declared in the header.
void ImputationResult::write_vcf(const std::string& filename, const TargetData& targets,
                                 const ReferencePanel& reference, const ImputationConfig& config) const {
    std::ofstream out(filename);
    if (!out) throw ImputationError("cannot open output VCF: " + filename);
    const bool ds = has_posteriors_ && config.output_dosages;
    const bool gp = has_posteriors_ && config.output_probabilities;
    const bool info = config.output_info_score;
    out << "##fileformat=VCFv4.2\n##source=SwiftImpute\n";
    if (!reference.markers().empty()) out << "##contig=<ID=" << reference.markers()[0].chrom << ">\n";
    out << "##INFO=<ID=IMP,Number=0,Type=Flag,Description=\"Imputed variant\">\n";
    if (info)
        out << "##INFO=<ID=INFO,Number=1,Type=Float,Description=\"Imputation quality: 1 - Var(DS) / "
               "(2p(1-p)), DS the "
            << (has_posteriors_ ? "posterior" : "sampled") << " dosage\">\n";
    out << "##FORMAT=<ID=GT,Number=1,Type=String,Description=\"Genotype\">\n";
    if (ds) out << "##FORMAT=<ID=DS,Number=1,Type=Float,Description=\"Dosage (expected ALT allele count)\">\n";
    if (gp) {
        out << "##FORMAT=<ID=GP,Number=G,Type=Float,Description=\"Genotype posterior probabilities\">\n";
        out << "##FORMAT=<ID=AP,Number=2,Type=Float,Description=\"Allelic probabilities\">\n";
    }
    out << "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT";
    for (const auto& s : targets.samples()) out << "\t" << s.id;
    out << "\n";
    std::string fmt = "GT";
    if (ds) fmt += ":DS";
    if (gp) fmt += ":GP:AP";
    char buf[96];
    for (marker_t m = 0; m < num_markers_; ++m) {
        const auto& mk = reference.markers()[m];
        out << mk.chrom << "\t" << mk.pos << "\t" << (mk.id.empty() ? "." : mk.id) << "\t" << mk.ref << "\t"
            << mk.alt << "\t.\tPASS\tIMP";
        if (info) {
            std::snprintf(buf, sizeof(buf), ";INFO=%.4f", double(info_scores_[m]));
            out << buf;
        }
        out << "\t" << fmt;
        for (sample_t s = 0; s < num_samples_; ++s) {
            const size_t b = size_t(s) * 2 * num_markers_;
            const allele_t a0 = haplotypes_[b + m], a1 = haplotypes_[b + num_markers_ + m];
            out << "\t";
            if (a0 == 255) out << ".";
            else out << int(a0);
            out << "|";
            if (a1 == 255) out << ".";
            else out << int(a1);
            if (ds) {
                std::snprintf(buf, sizeof(buf), ":%.4f", double(get_posterior_dosage(s, m)));
                out << buf;
            }
            if (gp) {
                prob_t g[3];
                get_probabilities(s, m, g);
                const size_t i = (size_t(s) * num_markers_ + m) * 2;
                std::snprintf(buf, sizeof(buf), ":%.4f,%.4f,%.4f:%.4f,%.4f", double(g[0]), double(g[1]),
                              double(g[2]), double(posteriors_[i]), double(posteriors_[i + 1]));
                out << buf;
            }
        }
        out << "\n";
    }
}

const allele_t* ImputationResult::haplotype0(sample_t s) const {
    return s < num_samples_ ? &haplotypes_[size_t(s) * 2 * num_markers_] : nullptr;
}

const allele_t* ImputationResult::haplotype1(sample_t s) const {
    return s < num_samples_ ? &haplotypes_[size_t(s) * 2 * num_markers_ + num_markers_] : nullptr;
}

prob_t ImputationResult::get_dosage(sample_t s, marker_t m) const {
    if (s >= num_samples_ || m >= num_markers_) return 0.0f;
    const size_t b = size_t(s) * 2 * num_markers_;
    return static_cast<prob_t>(haplotypes_[b + m] + haplotypes_[b + num_markers_ + m]);
}

void ImputationResult::get_probabilities(sample_t s, marker_t m, prob_t probs[3]) const {
    if (s >= num_samples_ || m >= num_markers_) {
        probs[0] = probs[1] = probs[2] = 0.0f;
        return;
    }
    const size_t i = (size_t(s) * num_markers_ + m) * 2;
    const prob_t p0 = posteriors_[i], p1 = posteriors_[i + 1];
    probs[0] = (1.0f - p0) * (1.0f - p1);
    probs[1] = (1.0f - p0) * p1 + p0 * (1.0f - p1);
    probs[2] = p0 * p1;
}

prob_t ImputationResult::get_posterior_dosage(sample_t s, marker_t m) const {
    if (s >= num_samples_ || m >= num_markers_) return 0.0f;
    const size_t i = (size_t(s) * num_markers_ + m) * 2;
    return posteriors_[i] + posteriors_[i + 1];
}

prob_t ImputationResult::get_info_score(marker_t m) const { return m < num_markers_ ? info_scores_[m] : 0.0f; }

void ImputationResult::set_haplotype(sample_t s, uint32_t phase, const allele_t* h) {
    if (s >= num_samples_ || phase >= 2) return;
    std::copy(h, h + num_markers_, &haplotypes_[size_t(s) * 2 * num_markers_ + size_t(phase) * num_markers_]);
}

void ImputationResult::set_posterior(sample_t s, marker_t m, const prob_t* p) {
    if (s >= num_samples_ || m >= num_markers_) return;
    const size_t i = (size_t(s) * num_markers_ + m) * 2;
    posteriors_[i] = p[0];
    posteriors_[i + 1] = p[1];
    has_posteriors_ = true;
}

// INFO = 1 - Var(d) / (2p(1-p)), p = mean(d) / 2, clamped to [0, 1], 1 for a
// monomorphic marker (src/api/imputer.cpp:467-496).  d is the posterior
// dosage AP0 + AP1 when the allele probabilities are filled, else (as the
// reference, whose posteriors stay empty) the sampled dosage h0 + h1.
void ImputationResult::compute_info_scores() {
    for (marker_t m = 0; m < num_markers_; ++m) {
        double sum = 0.0, sq = 0.0;
        for (sample_t s = 0; s < num_samples_; ++s) {
            const double d = has_posteriors_ ? double(get_posterior_dosage(s, m)) : double(get_dosage(s, m));
            sum += d;
            sq += d * d;
        }
        const double mean = sum / num_samples_, p = mean / 2.0;
        if (p == 0.0 || p == 1.0) {
            info_scores_[m] = 1.0f;
            continue;
        }
        const double var = sq / num_samples_ - mean * mean, ev = 2.0 * p * (1.0 - p);
        info_scores_[m] = static_cast<prob_t>(std::max(0.0, std::min(1.0, 1.0 - var / ev)));
    }
}

size_t ImputationResult::memory_usage() const {
    return size_t(num_samples_) * 2 * num_markers_ + size_t(num_samples_) * num_markers_ * 2 * sizeof(prob_t) +
           num_markers_ * sizeof(prob_t);
}

}  // namespace swiftimpute
