// Imputer: the reference's batch driver (src/api/imputer.cpp:547-1042) over
// the C ABI of include/swiftimpute_hip.h.
//
// build_index() keeps the bit-packed panel, the PBWT arrays and the
// sample-independent state table resident on the device.  impute() converts
// nothing per sample on the host beyond the palette coding TargetData already
// holds; per batch it uploads the codes, runs si_fb_run (one job per launch,
// the plain form) and copies haplotypes, allele probabilities and
// log-likelihoods back.  Without build_index(), selection falls back to the
// reference's simple pairs (2k, 2k+1) (imputer.cpp:929-942).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/swiftimpute/impute_c.h"
#include "../../../include/swiftimpute/imputer.hpp"
#include "../../../include/swiftimpute_hip.h"

namespace swiftimpute {

namespace {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e), int(e));
}

void si_check(int rc, const char* what) {
    if (rc == SI_OK) return;
    const std::string msg = std::string(what) + ": " + si_last_error();
    if (rc == SI_ERR_DEVICE) throw DeviceError(msg, rc);
    if (rc == SI_ERR_UNSUPPORTED) throw UnsupportedError(msg);
    throw ImputationError(msg);
}

template <class T>
T* dev_alloc(size_t count, size_t& total) {
    if (count == 0) return nullptr;
    void* p = nullptr;
    hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    total += count * sizeof(T);
    return static_cast<T*>(p);
}

// Frees p (allocated by dev_alloc with `count` elements) and takes its bytes
// off the running total.
template <class T>
void dev_free(T*& p, size_t count, size_t& total) {
    if (p) {
        (void)hipFree(p);
        total -= std::min(total, count * sizeof(T));
    }
    p = nullptr;
}

// Destructor-time release (no accounting).
template <class T>
void dev_drop(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Device memory a launch may take for its samples (checkpoints, codes,
// outputs).  Every draw depends only on (seed, sample, marker), so results do
// not depend on how samples are grouped into launches; one launch per cohort
// (up to this budget) keeps the chain pass, which takes one wavefront's walk
// over all markers however many samples run beside it, from repeating per
// batch_size samples.
constexpr size_t kLaunchBudget = size_t(16) << 30;

// Samples per launch: a multiple of batch_size (so progress is reported at
// the same sample counts as one launch per batch) within kLaunchBudget.
uint32_t launch_samples(uint32_t batch, uint32_t B, uint32_t M, uint32_t L, uint32_t window, uint32_t pitch) {
    si_fb_job j{};
    j.n_samples = 1;
    j.n_markers = M;
    j.n_states = L;
    j.window = window;
    const size_t per = si_fb_workspace_bytes(&j, 1) + pitch + size_t(M) * (2 * 4 + 2) + 8;
    const size_t groups = std::max<size_t>(1, kLaunchBudget / (per * batch + 1));
    const size_t want = std::min<size_t>(size_t(batch) * groups, (size_t(B) + batch - 1) / batch * batch);
    return uint32_t(std::max<size_t>(batch, std::min<size_t>(want, 0x7fffffffu)));
}

}  // namespace

struct Imputer::Device {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t bytes = 0;
    uint64_t* cols = nullptr;
    uint64_t* miss = nullptr;
    uint32_t* prefix = nullptr;
    uint32_t* div = nullptr;
    uint32_t* states = nullptr;
    uint8_t* sa = nullptr;
    uint8_t* mmask = nullptr;
    int table_mode = -1;  // 0 PBWT, 1 simple
    // per-batch
    uint32_t cap = 0;
    uint32_t pitch = 0;
    uint32_t cap_markers = 0;    // markers the per-batch buffers were sized for
    uint8_t* codes = nullptr;
    float* ap = nullptr;
    uint8_t* hap = nullptr;
    double* loglik = nullptr;
    void* ws = nullptr;
    size_t ws_bytes = 0;
    bool tuned = false;          // variant picked by si_fb_tune for this panel
    si_fb_variant variant{};
    si_step_coef* steps = nullptr;
    float* eps = nullptr;
    float* emax = nullptr;
    float* palette = nullptr;

    uint32_t n_palette = 0;      // palette entries of eps/emax/palette

    ~Device() {
        dev_drop(cols);
        dev_drop(miss);
        dev_drop(prefix);
        dev_drop(div);
        dev_drop(states);
        dev_drop(sa);
        dev_drop(mmask);
        dev_drop(codes);
        dev_drop(ap);
        dev_drop(hap);
        dev_drop(loglik);
        if (ws) (void)hipFree(ws);
        dev_drop(steps);
        dev_drop(eps);
        dev_drop(emax);
        dev_drop(palette);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

Imputer::Imputer(const ReferencePanel& reference, const ImputationConfig& config)
    : reference_(reference), config_(config), dev_(std::make_unique<Device>()) {
    const uint32_t L = config_.hmm_params.num_states;
    if (L == 0 || L > 256) throw ImputationError("num_states must be in 1..256");
    if (reference_.num_haplotypes() == 0) throw ImputationError("reference panel has no haplotypes");
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    dev_->device = config_.device_id < 0 ? 0 : config_.device_id;
    if (dev_->device >= n) throw ImputationError("device_id " + std::to_string(dev_->device) + " not present");
    hip_check(hipSetDevice(dev_->device), "hipSetDevice");
    hip_check(hipStreamCreate(&dev_->stream), "hipStreamCreate");
    const auto& c = reference_.packed_alleles();
    dev_->cols = dev_alloc<uint64_t>(c.size(), dev_->bytes);
    hip_check(hipMemcpyAsync(dev_->cols, c.data(), c.size() * 8, hipMemcpyHostToDevice, dev_->stream),
              "panel upload");
    const auto& ms = reference_.packed_missing();
    if (!ms.empty()) {
        dev_->miss = dev_alloc<uint64_t>(ms.size(), dev_->bytes);
        hip_check(hipMemcpyAsync(dev_->miss, ms.data(), ms.size() * 8, hipMemcpyHostToDevice, dev_->stream),
                  "panel upload");
    }
    hip_check(hipStreamSynchronize(dev_->stream), "panel upload");
}

Imputer::~Imputer() = default;

void Imputer::build_index() {
    hip_check(hipSetDevice(dev_->device), "hipSetDevice");
    const uint32_t M = reference_.num_markers(), H = reference_.num_haplotypes();
    const uint32_t L = config_.hmm_params.num_states, Lp = si_padded_states(L);
    if (!dev_->sa) {
        dev_->states = dev_alloc<uint32_t>(size_t(M) * L, dev_->bytes);
        dev_->sa = dev_alloc<uint8_t>(size_t(M) * Lp, dev_->bytes);
        dev_->mmask = dev_alloc<uint8_t>(M, dev_->bytes);
    }
    const int mode = config_.hmm_params.use_pbwt_selection ? 0 : 1;
    std::vector<uint32_t> perm(std::min(L, H));
    if (mode == 0) {
        // a rebuild reuses arrays a previous build kept
        if (!dev_->prefix) dev_->prefix = dev_alloc<uint32_t>(size_t(M) * H, dev_->bytes);
        if (!dev_->div) dev_->div = dev_alloc<uint32_t>(size_t(M) * H, dev_->bytes);
        const size_t wsb = si_pbwt_workspace_bytes(H, M);
        void* ws = nullptr;
        if (wsb) hip_check(hipMalloc(&ws, wsb), "hipMalloc");
        si_check(si_pbwt_sweep(dev_->cols, H, M, dev_->prefix, dev_->div, ws, wsb, dev_->stream),
                 "PBWT sweep");
        si_check(config_.selection == SelectionMode::GpuCompat ? si_select_perm_gpu_compat(H, L, perm.data())
                                                               : si_select_perm(H, L, perm.data()),
                 "selection permutation");
        si_check(si_state_table(dev_->prefix, dev_->cols, dev_->miss, H, M, L, 0, perm.data(), dev_->states,
                                dev_->sa, dev_->mmask, dev_->stream),
                 "state table");
        hip_check(hipStreamSynchronize(dev_->stream), "build_index");
        if (ws) (void)hipFree(ws);
        if (!config_.keep_pbwt_arrays) {
            dev_free(dev_->prefix, size_t(M) * H, dev_->bytes);
            dev_free(dev_->div, size_t(M) * H, dev_->bytes);
        }
    } else {
        si_check(si_state_table(nullptr, dev_->cols, dev_->miss, H, M, L, 1, perm.data(), dev_->states, dev_->sa,
                                dev_->mmask, dev_->stream),
                 "state table");
        hip_check(hipStreamSynchronize(dev_->stream), "build_index");
    }
    dev_->table_mode = mode;
    index_built_ = true;
    Logger::instance().info("index built: " + std::to_string(M) + " markers, " + std::to_string(H) +
                            " haplotypes, L=" + std::to_string(L));
}

void Imputer::validate_targets(const TargetData& targets) const {
    if (targets.num_markers() != reference_.num_markers())
        throw ImputationError("Target and reference marker counts don't match: " +
                              std::to_string(targets.num_markers()) + " vs " +
                              std::to_string(reference_.num_markers()));
}

std::unique_ptr<ImputationResult> Imputer::impute(const TargetData& targets) {
    return impute_with_progress(targets, nullptr);
}

std::unique_ptr<ImputationResult> Imputer::impute_with_progress(
    const TargetData& targets, std::function<void(uint32_t, uint32_t)> progress_callback) {
    validate_targets(targets);
    hip_check(hipSetDevice(dev_->device), "hipSetDevice");
    const uint32_t M = targets.num_markers(), B = targets.num_samples(), H = reference_.num_haplotypes();
    const uint32_t L = config_.hmm_params.num_states;
    if (dev_->table_mode < 0) {
        // No index: the reference falls back to simple selection (2k, 2k+1).
        if (2ull * L > H) throw ImputationError("simple state selection needs 2L <= haplotypes");
        const bool pbwt = config_.hmm_params.use_pbwt_selection;
        config_.hmm_params.use_pbwt_selection = false;
        build_index();
        config_.hmm_params.use_pbwt_selection = pbwt;
        index_built_ = false;
    }
    auto result = std::make_unique<ImputationResult>(B, M);
    if (B == 0 || M == 0) return result;

    // transitions from the reference panel's map (imputer.cpp:596-616)
    std::vector<uint64_t> pos(M);
    std::vector<double> cm(M), g(M);
    for (uint32_t m = 0; m < M; ++m) {
        pos[m] = reference_.markers()[m].pos;
        cm[m] = reference_.markers()[m].cM;
    }
    si_check(si_genetic_map(pos.data(), cm.data(), M, g.data()), "genetic map");
    std::vector<float> Td(std::max<uint32_t>(M, 2) - 1), To(Td.size());
    si_check(si_transition_constants(g.data(), M, config_.hmm_params.ne, config_.hmm_params.rho_rate, Td.data(),
                                     To.data()),
             "transitions");
    const uint32_t nint = M >= 2 ? M - 1 : 0;
    std::vector<si_step_coef> co(std::max<uint32_t>(nint, 1));
    si_check(si_step_coefs(Td.data(), To.data(), nint, co.data()), "step coefficients");
    const bool per_marker = config_.transition_mode == TransitionMode::PerMarker;
    si_step_coef step = nint ? co[0] : si_step_coef{1.0f, 0.0f, 0.0f, -INFINITY};
    // The linear kernel needs every switch probability it uses to be a normal
    // FP32 number; coincident markers (To = -999) need the log-space kernel.
    float min_switch = 1.0f;
    for (uint32_t t = 0; t < (per_marker ? nint : std::min<uint32_t>(nint, 1)); ++t)
        min_switch = std::min(min_switch, co[t].t_o);
    bool logspace = config_.precision == Precision::LogF32 ||
                    (config_.precision == Precision::Auto && !(min_switch >= 1e-30f));
    if (!logspace && !(min_switch >= 1e-30f))
        throw ImputationError("a switch probability is below the linear FP32 kernel's range "
                              "(coincident markers); use Precision::LogF32 or Precision::Auto");
    if (per_marker) {
        if (!dev_->steps) dev_->steps = dev_alloc<si_step_coef>(M, dev_->bytes);
        std::vector<si_step_coef> full(M, si_step_coef{1.0f, 0.0f, 0.0f, -INFINITY});
        for (uint32_t m = 1; m < M; ++m) full[m] = co[m - 1];
        hip_check(hipMemcpyAsync(dev_->steps, full.data(), M * sizeof(si_step_coef), hipMemcpyHostToDevice,
                                 dev_->stream),
                  "transition upload");
    }

    // emission table for this cohort's likelihood palette
    const auto& pal = targets.palette();
    const uint32_t P = uint32_t(pal.size());
    std::vector<float> palf(3 * P), eps(size_t(P) * 32), emax(size_t(P) * 8);
    for (uint32_t p = 0; p < P; ++p) {
        palf[3 * p] = pal[p].ll_00;
        palf[3 * p + 1] = pal[p].ll_01;
        palf[3 * p + 2] = pal[p].ll_11;
    }
    si_check(si_emission_table(palf.data(), P, eps.data(), emax.data()), "emission table");
    if (dev_->n_palette != P) {  // tables sized for this palette, reused across calls
        dev_free(dev_->eps, size_t(dev_->n_palette) * 32, dev_->bytes);
        dev_free(dev_->emax, size_t(dev_->n_palette) * 8, dev_->bytes);
        dev_free(dev_->palette, size_t(dev_->n_palette) * 3, dev_->bytes);
        dev_->eps = dev_alloc<float>(eps.size(), dev_->bytes);
        dev_->emax = dev_alloc<float>(emax.size(), dev_->bytes);
        dev_->palette = dev_alloc<float>(palf.size(), dev_->bytes);
        dev_->n_palette = P;
    }
    hip_check(hipMemcpyAsync(dev_->palette, palf.data(), palf.size() * 4, hipMemcpyHostToDevice,
                             dev_->stream),
              "upload");
    hip_check(hipMemcpyAsync(dev_->eps, eps.data(), eps.size() * 4, hipMemcpyHostToDevice, dev_->stream), "upload");
    hip_check(hipMemcpyAsync(dev_->emax, emax.data(), emax.size() * 4, hipMemcpyHostToDevice, dev_->stream),
              "upload");

    const uint32_t pitch = targets.code_pitch();
    const uint32_t report = std::max<uint32_t>(1, std::min(config_.batch_size, B));  // progress step
    const uint32_t batch = std::min(B, launch_samples(report, B, M, L, config_.window_size, pitch));
    if (dev_->cap < batch || dev_->pitch != pitch || dev_->cap_markers != M) {
        dev_free(dev_->codes, size_t(dev_->cap) * dev_->pitch, dev_->bytes);
        dev_free(dev_->ap, size_t(dev_->cap) * dev_->cap_markers * 2, dev_->bytes);
        dev_free(dev_->hap, size_t(dev_->cap) * 2 * dev_->cap_markers, dev_->bytes);
        dev_free(dev_->loglik, size_t(dev_->cap), dev_->bytes);
        dev_->codes = dev_alloc<uint8_t>(size_t(batch) * pitch, dev_->bytes);
        dev_->ap = dev_alloc<float>(size_t(batch) * M * 2, dev_->bytes);
        dev_->hap = dev_alloc<uint8_t>(size_t(batch) * 2 * M, dev_->bytes);
        dev_->loglik = dev_alloc<double>(batch, dev_->bytes);
        dev_->cap = batch;
        dev_->pitch = pitch;
        dev_->cap_markers = M;
    }
    si_fb_job job{};
    job.n_markers = M;
    job.n_states = L;
    job.flags = (per_marker ? uint32_t(SI_FB_PER_MARKER) : 0u) |
                (config_.sampling == SamplingMode::Argmax ? uint32_t(SI_FB_ARGMAX) : 0u) |
                (logspace ? uint32_t(SI_FB_LOGSPACE) : 0u) |
                // switch probabilities at the floor's scale: carry it exactly
                ((logspace && !(min_switch >= 0x1p-30f)) ? uint32_t(SI_FB_BANDED) : 0u);
    job.code_pitch = pitch;
    job.n_palette = P;
    for (uint32_t p = 0; p < P; ++p)  // untyped markers: likelihoods (0, 0, 0)
        if (palf[3 * p] == 0.0f && palf[3 * p + 1] == 0.0f && palf[3 * p + 2] == 0.0f) {
            job.flags |= SI_FB_UNTYPED;
            job.untyped_code = p;
            break;
        }
    job.seed = config_.seed;
    job.step = step;
    job.steps = dev_->steps;
    job.sa = dev_->sa;
    job.mmask = dev_->mmask;
    job.codes = dev_->codes;
    job.eps = dev_->eps;
    job.emax = dev_->emax;
    job.palette = dev_->palette;
    const bool want_ap = config_.output_probabilities || config_.output_dosages;
    job.ap = want_ap ? dev_->ap : nullptr;
    job.hap = dev_->hap;
    job.loglik = dev_->loglik;
    job.n_samples = batch;
    job.window = config_.window_size;
    const size_t need = si_fb_workspace_bytes(&job, 1);
    if (need > dev_->ws_bytes) {
        if (dev_->ws) (void)hipFree(dev_->ws);
        dev_->bytes -= dev_->ws_bytes;
        hip_check(hipMalloc(&dev_->ws, need), "hipMalloc");
        dev_->ws_bytes = need;
        dev_->bytes += need;
    }
    if (job.window && M > job.window)
        Logger::instance().info("windowed forward-backward (" + std::to_string(M) + " markers, window=" +
                                std::to_string(job.window) + ")");
    const uint8_t* codes = targets.codes().data();
    uint32_t reported = 0;
    for (uint32_t start = 0; start < B; start += batch) {
        const uint32_t end = std::min(start + batch, B), n = end - start;
        hip_check(hipMemcpyAsync(dev_->codes, codes + size_t(start) * pitch, size_t(n) * pitch,
                                 hipMemcpyHostToDevice, dev_->stream),
                  "codes upload");
        job.n_samples = n;
        job.sample_base = start;
        if (!dev_->tuned) {  // lane shapes and checkpoint block for this panel's jobs (cached per process)
            si_check(si_fb_tune(&job, 1, dev_->ws, dev_->ws_bytes, 0, dev_->stream, &dev_->variant),
                     "variant tuning");
            dev_->tuned = true;
        }
        si_check(si_fb_run_variant(&job, 1, &dev_->variant, dev_->ws, dev_->ws_bytes,
                                   SI_FB_PASS_CHAIN | SI_FB_PASS_BLOCK, dev_->stream),
                 "forward-backward");
        hip_check(hipMemcpyAsync(result->mutable_haplotypes() + size_t(start) * 2 * M, dev_->hap,
                                 size_t(n) * 2 * M, hipMemcpyDeviceToHost, dev_->stream),
                  "haplotype download");
        if (want_ap)
            hip_check(hipMemcpyAsync(result->mutable_posteriors() + size_t(start) * M * 2, dev_->ap,
                                     size_t(n) * M * 2 * sizeof(float), hipMemcpyDeviceToHost, dev_->stream),
                      "posterior download");
        hip_check(hipMemcpyAsync(result->mutable_log_likelihoods().data() + start, dev_->loglik,
                                 size_t(n) * sizeof(double), hipMemcpyDeviceToHost, dev_->stream),
                  "log-likelihood download");
        hip_check(hipStreamSynchronize(dev_->stream), "batch");
        while (progress_callback && reported < end) {  // once per batch_size samples
            reported = std::min(reported + report, B);
            progress_callback(reported, B);
        }
    }
    result->set_has_posteriors(want_ap);
    if (config_.output_info_score) result->compute_info_scores();
    return result;
}

size_t Imputer::device_memory_usage() const { return dev_->bytes; }

size_t Imputer::estimate_memory_requirement(const ReferencePanel& reference, uint32_t num_target_samples,
                                            const ImputationConfig& config) {
    const size_t M = reference.num_markers(), H = reference.num_haplotypes();
    const uint32_t L = config.hmm_params.num_states, Lp = si_padded_states(L);
    const size_t W = (H + 63) / 64;
    size_t total = M * W * 8 * (reference.packed_missing().empty() ? 1 : 2);
    if (config.hmm_params.use_pbwt_selection) total += 2 * M * H * 4;
    total += M * (size_t(L) * 4 + Lp + 1);
    const uint32_t pitch = (uint32_t(M) + 15) / 16 * 16;
    const uint32_t report = std::max<uint32_t>(1, std::min(config.batch_size, num_target_samples));
    const uint32_t batch = std::max<uint32_t>(
        1, std::min(num_target_samples,
                    launch_samples(report, num_target_samples, uint32_t(M), L, config.window_size, pitch)));
    si_fb_job j{};
    j.n_samples = batch;
    j.n_markers = uint32_t(M);
    j.n_states = L;
    j.window = config.window_size;
    total += si_fb_workspace_bytes(&j, 1);
    total += size_t(batch) * ((M + 15) / 16 * 16 + M * 2 * 4 + 2 * M + 8);
    return total;
}

std::vector<uint32_t> Imputer::pbwt_prefix() const {
    std::vector<uint32_t> out;
    if (!dev_->prefix) return out;
    out.resize(size_t(reference_.num_markers()) * reference_.num_haplotypes());
    hip_check(hipMemcpy(out.data(), dev_->prefix, out.size() * 4, hipMemcpyDeviceToHost), "prefix download");
    return out;
}

std::vector<uint32_t> Imputer::pbwt_divergence() const {
    std::vector<uint32_t> out;
    if (!dev_->div) return out;
    out.resize(size_t(reference_.num_markers()) * reference_.num_haplotypes());
    hip_check(hipMemcpy(out.data(), dev_->div, out.size() * 4, hipMemcpyDeviceToHost), "divergence download");
    return out;
}

std::vector<uint32_t> Imputer::selected_states() const {
    std::vector<uint32_t> out;
    if (!dev_->states) return out;
    out.resize(size_t(reference_.num_markers()) * config_.hmm_params.num_states);
    hip_check(hipMemcpy(out.data(), dev_->states, out.size() * 4, hipMemcpyDeviceToHost), "states download");
    return out;
}

std::unique_ptr<ImputationResult> impute(const std::string& reference_vcf, const std::string& target_vcf,
                                         const std::string& output_vcf, const ImputationConfig& config) {
    auto reference = ReferencePanel::load_vcf(reference_vcf);
    auto targets = TargetData::load_vcf(target_vcf);
    Imputer imputer(*reference, config);
    try {
        imputer.build_index();
    } catch (const UnsupportedError&) {
        throw;  // a panel this build cannot index is an error, not a fallback
    } catch (const ImputationError& e) {
        Logger::instance().warning(std::string("PBWT index not available, proceeding without it: ") + e.what());
    }
    auto result = imputer.impute(*targets);
    result->write_vcf(output_vcf, *targets, *reference, config);
    return result;
}

}  // namespace swiftimpute

// ------------------------------------------------------------ C entry point --
namespace {
thread_local std::string g_impute_err;
}

extern "C" {

void si_impute_defaults(si_impute_opts* o) {
    if (!o) return;
    const swiftimpute::ImputationConfig c;
    o->num_states = c.hmm_params.num_states;
    o->ne = c.hmm_params.ne;
    o->batch_size = c.batch_size;
    o->window_size = c.window_size;
    o->precision = 0;
    o->per_marker = 0;
    o->seed = c.seed;
}

const char* si_impute_last_error(void) { return g_impute_err.c_str(); }

int si_impute_vcf(const char* reference_vcf, const char* target_vcf, const char* output_vcf,
                  const si_impute_opts* opts, si_impute_timing* t) {
    using namespace swiftimpute;
    using Clock = std::chrono::steady_clock;
    auto secs = [](Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    if (!reference_vcf || !target_vcf || !output_vcf || !opts || opts->precision > 2) {
        g_impute_err = "si_impute_vcf: bad arguments";
        return SI_ERR_ARG;
    }
    try {
        ImputationConfig cfg;
        cfg.hmm_params.num_states = opts->num_states;
        cfg.hmm_params.ne = opts->ne;
        cfg.batch_size = opts->batch_size;
        cfg.window_size = opts->window_size;
        cfg.precision = static_cast<Precision>(opts->precision);
        cfg.transition_mode = opts->per_marker ? TransitionMode::PerMarker : TransitionMode::FirstInterval;
        cfg.seed = opts->seed;
        const auto t0 = Clock::now();
        auto reference = ReferencePanel::load_vcf(reference_vcf);
        const auto t1 = Clock::now();
        auto targets = TargetData::load_vcf(target_vcf);
        const auto t2 = Clock::now();
        Imputer imputer(*reference, cfg);
        try {
            imputer.build_index();
        } catch (const UnsupportedError&) {
            throw;
        } catch (const DeviceError&) {
            throw;
        } catch (const ImputationError& e) {  // as the reference: proceed without the index
            Logger::instance().warning(std::string("PBWT index not available, proceeding without it: ") + e.what());
        }
        const auto t3 = Clock::now();
        auto result = imputer.impute(*targets);
        const auto t4 = Clock::now();
        result->write_vcf(output_vcf, *targets, *reference, cfg);
        const auto t5 = Clock::now();
        if (t) {
            t->reference_load_s = secs(t0, t1);
            t->target_load_s = secs(t1, t2);
            t->index_s = secs(t2, t3);
            t->impute_s = secs(t3, t4);
            t->write_s = secs(t4, t5);
            t->total_s = secs(t0, t5);
            t->samples = targets->num_samples();
            t->markers = targets->num_markers();
            uint64_t untyped = 0;
            const auto& pal = targets->palette();
            const auto& codes = targets->codes();
            for (size_t s = 0; s < targets->num_samples(); ++s)
                for (size_t m = 0; m < targets->num_markers(); ++m) {
                    const auto& g = pal[codes[s * targets->code_pitch() + m]];
                    untyped += g.ll_00 == 0.0f && g.ll_01 == 0.0f && g.ll_11 == 0.0f;
                }
            t->untyped_genotypes = untyped;
        }
        return SI_OK;
    } catch (const UnsupportedError& e) {
        g_impute_err = e.what();
        return SI_ERR_UNSUPPORTED;
    } catch (const DeviceError& e) {
        g_impute_err = e.what();
        return SI_ERR_DEVICE;
    } catch (const std::exception& e) {
        g_impute_err = e.what();
        return SI_ERR_ARG;
    }
}

}  // extern "C"
