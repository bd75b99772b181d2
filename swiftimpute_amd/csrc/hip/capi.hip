// extern "C" entry points of libsi_hip.so (declared in include/swiftimpute_hip.h).
//
// Argument checking, host-side constant preparation (selection permutation,
// transition and emission tables) and kernel dispatch.  Errors are returned as
// status codes with a thread-local message; nothing here throws across the
// C boundary.
#include <algorithm>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "si_internal.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int device_status(int rc, const char* what) {
    if (rc == 0) return SI_OK;
    if (rc == SI_ERR_UNSUPPORTED) return fail(rc, "%s: shape not supported by this build", what);
    const hipError_t e = hipGetLastError();
    return fail(SI_ERR_DEVICE, "%s: HIP error %d (%s)", what, int(e), hipGetErrorString(e));
}

bool aligned(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; }

}  // namespace

extern "C" {

const char* si_last_error(void) { return g_last_error.c_str(); }

int si_abi_version(void) { return SI_ABI_VERSION; }

uint32_t si_padded_states(uint32_t L) {
    uint32_t lp = 4;
    while (lp < L) lp <<= 1;
    return lp;
}

// ------------------------------------------------------------------- PBWT --
size_t si_pbwt_workspace_bytes(uint32_t H, uint32_t M) {
    return si::pbwt_workspace_bytes(H, M, si::pbwt_default_chunk(H, M));
}

size_t si_pbwt_chunked_workspace_bytes(uint32_t H, uint32_t M, uint32_t chunk) {
    return si::pbwt_workspace_bytes(H, M, chunk);
}

int si_pbwt_sweep_chunked(const uint64_t* cols, uint32_t H, uint32_t M, uint32_t* prefix,
                          uint32_t* div, uint32_t chunk, void* workspace, size_t workspace_bytes,
                          void* stream) {
    if (H == 0) return fail(SI_ERR_ARG, "si_pbwt_sweep: H must be positive");
    if (H > si::kPbwtMaxH)
        return fail(SI_ERR_UNSUPPORTED, "si_pbwt_sweep: H = %u above the supported %u haplotypes", H,
                    si::kPbwtMaxH);
    if (chunk == 0) return fail(SI_ERR_ARG, "si_pbwt_sweep: chunk must be positive");
    if (M == 0) return SI_OK;
    if (!cols || !prefix) return fail(SI_ERR_ARG, "si_pbwt_sweep: cols and prefix are required");
    auto st = static_cast<hipStream_t>(stream);
    const size_t need = si::pbwt_workspace_bytes(H, M, chunk);
    void* ws = workspace;
    if (!ws) {  // stream-ordered scratch of the call's own
        if (hipMallocAsync(&ws, need, st) != hipSuccess)
            return fail(SI_ERR_DEVICE, "si_pbwt_sweep: workspace allocation of %zu bytes failed", need);
    } else if (workspace_bytes < need) {
        return fail(SI_ERR_WORKSPACE, "si_pbwt_sweep: workspace of %zu bytes, %zu needed",
                    workspace_bytes, need);
    }
    const int rc = si::launch_pbwt_sweep(cols, H, M, prefix, div, chunk, ws, need, st);
    if (!workspace) (void)hipFreeAsync(ws, st);
    return device_status(rc, "si_pbwt_sweep");
}

int si_pbwt_sweep(const uint64_t* cols, uint32_t H, uint32_t M, uint32_t* prefix, uint32_t* div,
                  void* workspace, size_t workspace_bytes, void* stream) {
    return si_pbwt_sweep_chunked(cols, H, M, prefix, div, si::pbwt_default_chunk(H, M), workspace,
                                 workspace_bytes, stream);
}

// -------------------------------------------------------------- selection --
int si_select_perm(uint32_t H, uint32_t L, uint32_t* perm) {
    if (H == 0 || L == 0 || !perm) return fail(SI_ERR_ARG, "si_select_perm: bad arguments");
    // The reference sorts (divergence, haplotype) candidates with this exact
    // comparator (src/pbwt/pbwt_index.cpp:30-33); all divergences in a row are
    // equal, so the result is a fixed permutation of the prefix row.
    std::vector<std::pair<uint32_t, uint32_t>> cand(H);
    for (uint32_t i = 0; i < H; ++i) cand[i] = {0u, i};
    std::sort(cand.begin(), cand.end(),
              [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
                  return a.first > b.first;
              });
    const uint32_t n = std::min(L, H);
    for (uint32_t k = 0; k < n; ++k) perm[k] = cand[k].second;
    return SI_OK;
}

int si_select_perm_gpu_compat(uint32_t H, uint32_t L, uint32_t* perm) {
    if (H == 0 || L == 0 || L > si::kMaxStates || !perm)
        return fail(SI_ERR_ARG, "si_select_perm_gpu_compat: need H > 0, 1 <= L <= %u, perm", si::kMaxStates);
    // launch_select_states (src/pbwt/pbwt_selector.cu:400-463) with every
    // divergence in a row equal.  Warp kernel (L <= 32, H <= 10000,
    // :212-312): each lane keeps its first four strided entries, every round's
    // maximum is tied and lane 0 wins it, so round r emits lane 0's r-th entry,
    // prefix position 32r, for r below lane 0's count min(4, ceil(H/32)); the
    // rest is filled with haplotype k (k % H, :305-310).  Insertion kernel
    // (:320-394, 32 < L <= 64 or H > 10000): ties never displace, so the
    // first min(L, H) positions in prefix order.  Block kernel (L > 64,
    // :47-204): thread t keeps the first min(L, n_t) positions of its range
    // [t*ceil(H/256), ...) in order (ties never displace), and the merge in
    // thread order appends them until L: again the first min(L, H) positions
    // (a thread whose range holds fewer than L leaves the rest of its slots
    // unwritten there and the merge reads them; taken here as empty).
    const uint32_t n = std::min(L, H);
    if (L <= 32 && H <= 10000) {
        const uint32_t c0 = std::min(4u, (H + 31) / 32);
        for (uint32_t k = 0; k < n; ++k) perm[k] = k < c0 ? 32 * k : (SI_PERM_FIXED | k);
    } else {
        for (uint32_t k = 0; k < n; ++k) perm[k] = k;
    }
    return SI_OK;
}

int si_state_table(const uint32_t* prefix, const uint64_t* cols, const uint64_t* miss, uint32_t H,
                   uint32_t M, uint32_t L, uint32_t mode, const uint32_t* perm_host,
                   uint32_t* states, uint8_t* sa, uint8_t* mmask, void* stream) {
    if (H == 0 || L == 0 || L > si::kMaxStates)
        return fail(SI_ERR_ARG, "si_state_table: need H > 0 and 1 <= L <= %u (got H=%u L=%u)", si::kMaxStates,
                    H, L);
    if (mode > 1) return fail(SI_ERR_ARG, "si_state_table: unknown mode %u", mode);
    if (mode == 1 && 2ull * L > H)
        return fail(SI_ERR_ARG, "si_state_table: simple selection needs 2L <= H");
    if (mode == 0 && (!prefix || !perm_host))
        return fail(SI_ERR_ARG, "si_state_table: PBWT selection needs prefix and perm");
    if (!cols || !sa || !mmask) return fail(SI_ERR_ARG, "si_state_table: cols, sa, mmask required");
    if (mode == 0)
        for (uint32_t k = 0; k < std::min(L, H); ++k)
            if ((perm_host[k] & ~SI_PERM_FIXED) >= H)
                return fail(SI_ERR_ARG, "si_state_table: perm[%u] = 0x%x is outside the panel (H=%u)",
                            k, perm_host[k], H);
    if (M == 0) return SI_OK;
    const uint32_t Lp = si_padded_states(L);
    return device_status(si::launch_state_table(prefix, cols, miss, H, M, L, Lp, mode, perm_host,
                                                states, sa, mmask, static_cast<hipStream_t>(stream)),
                         "si_state_table");
}

// ------------------------------------------------------------ transitions --
int si_genetic_map(const uint64_t* pos, const double* cM, uint32_t M, double* g) {
    if ((!pos && M) || (!g && M)) return fail(SI_ERR_ARG, "si_genetic_map: null pointer");
    for (uint32_t m = 0; m < M; ++m) {
        const double c = cM ? cM[m] : 0.0;
        g[m] = (c > 0.0) ? c / 100.0 : static_cast<double>(pos[m]) / 1e8;
    }
    return SI_OK;
}

int si_transition_constants(const double* g, uint32_t M, uint32_t ne, float rho_rate, float* Td,
                            float* To) {
    if (M >= 2 && (!g || !Td || !To)) return fail(SI_ERR_ARG, "si_transition_constants: null pointer");
    if (ne == 0) return fail(SI_ERR_ARG, "si_transition_constants: ne must be positive");
    const double dne = static_cast<double>(ne);
    const double rr = static_cast<double>(rho_rate);
    for (uint32_t t = 0; t + 1 < M; ++t) {
        double d = g[t + 1] - g[t];
        if (d < 0.0) d = 0.0;
        const double exponent = -4.0 * dne * d * rr;
        double rho = 1.0 - std::exp(exponent);
        if (rho < 0.0) rho = 0.0;
        if (rho > 1.0) rho = 1.0;
        Td[t] = static_cast<float>(std::log10((1.0 - rho) + rho / (2.0 * dne)));
        const double sw = rho / (2.0 * dne);
        To[t] = (sw < 1e-300) ? -999.0f : static_cast<float>(std::log10(sw));
    }
    return SI_OK;
}

int si_step_coefs(const float* Td, const float* To, uint32_t n, si_step_coef* out) {
    if (n && (!Td || !To || !out)) return fail(SI_ERR_ARG, "si_step_coefs: null pointer");
    for (uint32_t t = 0; t < n; ++t) {
        const double ed = std::exp(double(Td[t])), eo = std::exp(double(To[t]));
        out[t].t_o = static_cast<float>(eo);
        out[t].u = static_cast<float>(std::max(ed - eo, 0.0));
        out[t].lto = To[t];
        out[t].lu = (ed > eo) ? static_cast<float>(std::log(ed - eo)) : -INFINITY;
    }
    return SI_OK;
}

// --------------------------------------------------------------- emission --
int si_emission_table(const float* palette, uint32_t P, float* eps, float* emax) {
    if (!palette || !eps || !emax || P == 0) return fail(SI_ERR_ARG, "si_emission_table: bad arguments");
    for (uint32_t p = 0; p < P; ++p) {
        for (uint32_t c = 0; c < 8; ++c) {
            double mx = -INFINITY;
            for (int k = 0; k < 3; ++k)
                if (c & (1u << k)) mx = std::max(mx, double(palette[3 * p + k]));
            if (c == 0) mx = 0.0;
            float* row = eps + (size_t(p) * 8 + c) * 4;
            for (int k = 0; k < 3; ++k)
                row[k] = (c & (1u << k)) ? static_cast<float>(std::exp(double(palette[3 * p + k]) - mx))
                                         : 0.0f;
            row[3] = 0.0f;
            emax[size_t(p) * 8 + c] = static_cast<float>(mx);
        }
    }
    return SI_OK;
}

// -------------------------------------------------------- forward-backward --
static int check_job(const si_fb_job& j, uint32_t idx) {
    if (j.n_states == 0 || j.n_states > si::kMaxStates)
        return fail(SI_ERR_ARG, "job %u: n_states must be in 1..%u (got %u)", idx, si::kMaxStates, j.n_states);
    if (j.n_palette == 0 || j.n_palette > si::kMaxPalette)
        return fail(SI_ERR_UNSUPPORTED, "job %u: palette size %u outside 1..%u", idx, j.n_palette,
                    si::kMaxPalette);
    if (j.code_pitch < j.n_markers || (j.code_pitch & 15u))
        return fail(SI_ERR_ARG, "job %u: code_pitch must be >= M and a multiple of 16", idx);
    if (j.n_samples == 0 || j.n_markers == 0) return SI_OK;
    const bool logspace = (j.flags & SI_FB_LOGSPACE) != 0;
    if (!j.sa || !j.mmask || !j.codes)
        return fail(SI_ERR_ARG, "job %u: sa, mmask, codes are required", idx);
    if (!j.emax || (logspace ? !j.palette : !j.eps))
        return fail(SI_ERR_ARG, "job %u: %s required", idx,
                    logspace ? "palette and emax are" : "eps and emax are");
    if (!aligned(j.codes, 16) || !aligned(j.sa, 4) || (j.eps && !aligned(j.eps, 16)))
        return fail(SI_ERR_ARG, "job %u: codes (16 B), sa (4 B), eps (16 B) must be aligned", idx);
    if ((j.flags & SI_FB_PER_MARKER) && !j.steps)
        return fail(SI_ERR_ARG, "job %u: per-marker transitions need steps[]", idx);
    if (!logspace && !(j.flags & SI_FB_PER_MARKER) && j.n_markers >= 2 && !(j.step.t_o >= 1e-30f))
        return fail(SI_ERR_UNSUPPORTED,
                    "job %u: switch probability e^To = %g below the linear-space range "
                    "(use SI_FB_LOGSPACE)", idx, double(j.step.t_o));
    if (!j.ds && !j.ap && !j.hap && !j.prob && !j.loglik)
        return fail(SI_ERR_ARG, "job %u: no output requested", idx);
    if (si::fb_window(j) && j.window < si::kMinWindow)
        return fail(SI_ERR_UNSUPPORTED, "job %u: window %u shorter than %u markers", idx, j.window,
                    si::kMinWindow);
    if ((j.flags & SI_FB_UNTYPED) && j.untyped_code >= j.n_palette)
        return fail(SI_ERR_ARG, "job %u: untyped_code %u outside the palette (%u entries)", idx,
                    j.untyped_code, j.n_palette);
    return SI_OK;
}

size_t si_fb_workspace_bytes(const si_fb_job* jobs, uint32_t n_jobs) {
    size_t mx = 0;
    for (uint32_t i = 0; jobs && i < n_jobs; ++i)
        mx = std::max(mx, si::fb_workspace_bytes_one(jobs[i].n_samples, jobs[i].n_markers,
                                                     si_padded_states(jobs[i].n_states),
                                                     si::fb_window(jobs[i]), si::fb_band_job(jobs[i])));
    return mx;
}

int si_fb_default_variant(uint32_t n_states, si_fb_variant* out) {
    if (!out || n_states == 0 || n_states > si::kMaxStates)
        return fail(SI_ERR_ARG, "si_fb_default_variant: n_states must be in 1..%u", si::kMaxStates);
    const si::FbVariant v = si::fb_default_variant(si_padded_states(n_states));
    *out = si_fb_variant{v.gc, v.gb, v.cb};
    return SI_OK;
}

uint32_t si_fb_lane_candidates(uint32_t n_states, uint32_t pass, uint32_t* lanes, uint32_t max) {
    if (n_states == 0 || n_states > si::kMaxStates || (pass != SI_FB_PASS_CHAIN && pass != SI_FB_PASS_BLOCK))
        return 0;
    return si::fb_candidates(si_padded_states(n_states), pass == SI_FB_PASS_CHAIN ? 0 : 1, lanes,
                             lanes ? max : 0);
}

int si_fb_run(const si_fb_job* jobs, uint32_t n_jobs, void* workspace, size_t workspace_bytes,
              void* stream) {
    return si_fb_run_variant(jobs, n_jobs, nullptr, workspace, workspace_bytes,
                             SI_FB_PASS_CHAIN | SI_FB_PASS_BLOCK, stream);
}

int si_fb_run_passes(const si_fb_job* jobs, uint32_t n_jobs, void* workspace,
                     size_t workspace_bytes, uint32_t passes, void* stream) {
    return si_fb_run_variant(jobs, n_jobs, nullptr, workspace, workspace_bytes, passes, stream);
}

extern "C++" {
static int check_variant(uint32_t Lp, const si_fb_variant* v, const char* what) {
    if (!v) return SI_OK;
    if (!si::fb_variant_valid(Lp, si::FbVariant{v->chain_lanes, v->block_lanes, v->ckpt_block}))
        return fail(SI_ERR_UNSUPPORTED,
                    "%s: variant (chain %u lanes, block %u lanes, checkpoint block %u) not built "
                    "for %u padded states", what, v->chain_lanes, v->block_lanes, v->ckpt_block, Lp);
    return SI_OK;
}

static si::FbVariant variant_or_default(uint32_t Lp, const si_fb_variant* v) {
    return v ? si::FbVariant{v->chain_lanes, v->block_lanes, v->ckpt_block}
             : si::fb_default_variant(Lp);
}
}  // extern "C++"

int si_fb_run_variant(const si_fb_job* jobs, uint32_t n_jobs, const si_fb_variant* variant,
                      void* workspace, size_t workspace_bytes, uint32_t passes, void* stream) {
    if (n_jobs && !jobs) return fail(SI_ERR_ARG, "si_fb_run: jobs is null");
    if (passes == 0 || (passes & ~3u))
        return fail(SI_ERR_ARG, "si_fb_run_passes: passes must be a nonzero CHAIN|BLOCK mask");
    // Jobs share one workspace in stream order, so a split sequence
    // (all chains, then all blocks) is only meaningful for a single job.
    if (passes != 3u && n_jobs > 1)
        return fail(SI_ERR_ARG, "si_fb_run_passes: a single pass runs one job at a time");
    for (uint32_t i = 0; i < n_jobs; ++i) {
        int rc = check_job(jobs[i], i);
        if (!rc) rc = check_variant(si_padded_states(jobs[i].n_states), variant, "si_fb_run_variant");
        if (rc) return rc;
    }
    const size_t need = si_fb_workspace_bytes(jobs, n_jobs);
    if (need && (!workspace || workspace_bytes < need))
        return fail(SI_ERR_WORKSPACE, "si_fb_run: workspace of %zu bytes needed, %zu given", need,
                    workspace_bytes);
    for (uint32_t i = 0; i < n_jobs; ++i) {
        const si_fb_job& j = jobs[i];
        if (j.n_samples == 0 || j.n_markers == 0) continue;
        const uint32_t Lp = si_padded_states(j.n_states);
        const si::FbVariant var = variant_or_default(Lp, variant);
        si::JobView v = si::fb_job_view(j, var.cb);
        si::fb_bind_workspace(v, Lp, static_cast<char*>(workspace));
        const si::FbTables t{j.eps, j.emax, j.palette, j.n_palette, si::fb_untyped_code(j)};
        const int rc = si::launch_fb(v, t, Lp, passes, var, static_cast<hipStream_t>(stream));
        if (rc) return device_status(rc, "si_fb_run");
    }
    return SI_OK;
}

}  // extern "C"

// --------------------------------------------- forward-backward, no allocation --
// The checkpoint workspace of si_fb_run_noalloc: one array per device, resident
// from the library's load until the process ends.  In host code `scratch`
// names a host-side stand-in, not device memory: only the address returned
// by hipGetSymbolAddress for the current device may reach a kernel or a copy.
namespace {
static __device__ __align__(256) char scratch[SI_FB_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// This device's address of `scratch`, looked up once per device (with that
// device current) and kept.
int scratch_address(int dev, void** out) {
    if (dev < 0 || dev >= kMaxDevices)
        return fail(SI_ERR_UNSUPPORTED, "si_fb_run_noalloc: device %d outside 0..%d", dev,
                    kMaxDevices - 1);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SI_ERR_DEVICE, "si_fb_run_noalloc: hipGetSymbolAddress failed on device %d: %s",
                        dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SI_OK;
}
}  // namespace

extern "C" {

int si_fb_run_noalloc(const si_fb_job* jobs, uint32_t n_jobs, const si_fb_variant* variant,
                      void* stream) {
    if (n_jobs && !jobs) return fail(SI_ERR_ARG, "si_fb_run_noalloc: jobs is null");
    const size_t need = si_fb_workspace_bytes(jobs, n_jobs);
    if (need > SI_FB_NOALLOC_SCRATCH_BYTES)
        return fail(SI_ERR_WORKSPACE,
                    "si_fb_run_noalloc: these jobs need %zu bytes of checkpoint workspace, above the "
                    "%u-byte static scratch; use si_fb_run with a caller workspace",
                    need, unsigned(SI_FB_NOALLOC_SCRATCH_BYTES));
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return fail(SI_ERR_DEVICE, "si_fb_run_noalloc: hipGetDevice failed");
    void* ws = nullptr;
    const int rc = scratch_address(dev, &ws);
    if (rc) return rc;
    // every checkpoint the block pass reads is written by the chain pass of
    // the same call, so nothing needs clearing between calls
    return si_fb_run_variant(jobs, n_jobs, variant, ws, SI_FB_NOALLOC_SCRATCH_BYTES,
                             SI_FB_PASS_CHAIN | SI_FB_PASS_BLOCK, stream);
}

}  // extern "C"

// ------------------------------------------------------------- autotuner --
namespace {

// Checkpoint blocks the tuners consider: 16 markers only (32-marker blocks
// were dropped: their block kernels spilled to scratch).
constexpr int kNumCkpt = 1;
constexpr uint32_t kCkptBlocks[kNumCkpt] = {16};

// Event pair timing one launch sequence on a stream (milliseconds).
struct Timer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = false;
    Timer() { ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess; }
    ~Timer() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    template <class F>
    int time_once(hipStream_t st, F&& f, float* ms) {
        if (hipEventRecord(e0, st) != hipSuccess) return SI_ERR_DEVICE;
        const int rc = f();
        if (rc) return rc;
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(ms, e0, e1) != hipSuccess)
            return SI_ERR_DEVICE;
        return 0;
    }
    // Runs f twice (the first warms caches and code objects), times the second.
    template <class F>
    int time(hipStream_t st, F&& f, float* ms) {
        const int rc = f();
        return rc ? rc : time_once(st, f, ms);
    }
};

// A job cut to its first `cap` markers with every output disabled.
si::JobView tuning_view(const si_fb_job& j, uint32_t cap, uint32_t cb) {
    si_fb_job c = j;
    c.n_markers = std::min(j.n_markers, cap);
    c.ds = c.ap = c.prob = nullptr;
    c.hap = nullptr;
    c.loglik = nullptr;
    return si::fb_job_view(c, cb);
}

struct CacheKey {
    int device;
    uint32_t Lp, flags, pm, bucket;
    uint32_t side;  // 1: a class of a batch whose classes run side by side
    bool operator==(const CacheKey& o) const {
        return device == o.device && Lp == o.Lp && flags == o.flags && pm == o.pm &&
               bucket == o.bucket && side == o.side;
    }
};
std::mutex g_tune_mu;
std::vector<std::pair<CacheKey, si::FbVariant>> g_tune_cache;

uint32_t size_bucket(uint64_t chains) {
    uint32_t b = 0;
    while (chains >= 4 && b < 31) {
        chains >>= 2;
        ++b;
    }
    return b;
}

bool cache_get(const CacheKey& k, si::FbVariant* v) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& e : g_tune_cache)
        if (e.first == k) {
            *v = e.second;
            return true;
        }
    return false;
}

void cache_put(const CacheKey& k, const si::FbVariant& v) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& e : g_tune_cache)
        if (e.first == k) {
            e.second = v;
            return;
        }
    g_tune_cache.emplace_back(k, v);
}

// SI_FB_TUNE_LOG=1 prints every timing the tuners take to stderr.
bool tune_log() {
    static const bool on = [] {
        const char* e = std::getenv("SI_FB_TUNE_LOG");
        return e && *e && *e != '0';
    }();
    return on;
}

// Best variant from per-(checkpoint block, lanes) times of each pass.
// The variant minimising chain + block time.  side: the class runs beside
// other classes, where isolated timings carry over only in part (measured: a
// block shape 7 % faster alone was 2 % slower in the batch) -- the chain keeps
// the default lanes, the block takes another shape only if it is at least
// 10 % faster alone.
si::FbVariant pick(const float chain_ms[kNumCkpt][8], const uint32_t* gc, uint32_t nc,
                   const float block_ms[kNumCkpt][8], const uint32_t* gb, uint32_t nb, uint32_t Lp,
                   bool side = false) {
    if (tune_log())
        for (int k = 0; k < kNumCkpt; ++k) {
            for (uint32_t i = 0; i < nc; ++i)
                std::fprintf(stderr, "[si_fb_tune] Lp=%u ckpt=%u chain lanes=%u: %.3f ms\n", Lp,
                             kCkptBlocks[k], gc[i], double(chain_ms[k][i]));
            for (uint32_t i = 0; i < nb; ++i)
                std::fprintf(stderr, "[si_fb_tune] Lp=%u ckpt=%u block lanes=%u: %.3f ms\n", Lp,
                             kCkptBlocks[k], gb[i], double(block_ms[k][i]));
        }
    si::FbVariant best{gc[0], gb[0], kCkptBlocks[0]};
    float best_ms = 3.4e38f;
    const si::FbVariant def = si::fb_default_variant(Lp);
    for (int k = 0; k < kNumCkpt; ++k) {
        uint32_t ic = 0, ib = 0;
        for (uint32_t i = 1; i < nc; ++i)
            if (chain_ms[k][i] < chain_ms[k][ic]) ic = i;
        for (uint32_t i = 1; i < nb; ++i)
            if (block_ms[k][i] < block_ms[k][ib]) ib = i;
        if (side) {
            uint32_t dc = 0, db = 0;  // the default shapes' places among the candidates
            for (uint32_t i = 0; i < nc; ++i)
                if (gc[i] == def.gc) dc = i;
            for (uint32_t i = 0; i < nb; ++i)
                if (gb[i] == def.gb) db = i;
            ic = dc;
            if (!(block_ms[k][ib] < 0.9f * block_ms[k][db])) ib = db;
        }
        const float t = chain_ms[k][ic] + block_ms[k][ib];
        if (t < best_ms) {
            best_ms = t;
            best = si::FbVariant{gc[ic], gb[ib], kCkptBlocks[k]};
        }
    }
    return best;
}

}  // namespace

extern "C" {

int si_fb_tune(const si_fb_job* jobs, uint32_t n_jobs, void* workspace, size_t workspace_bytes,
               uint32_t max_markers, void* stream_v, si_fb_variant* out) {
    if (!jobs || n_jobs == 0 || !out) return fail(SI_ERR_ARG, "si_fb_tune: bad arguments");
    for (uint32_t i = 0; i < n_jobs; ++i) {
        const int rc = check_job(jobs[i], i);
        if (rc) return rc;
        if (si_padded_states(jobs[i].n_states) != si_padded_states(jobs[0].n_states) ||
            (jobs[i].flags & SI_FB_LOGSPACE) != (jobs[0].flags & SI_FB_LOGSPACE))
            return fail(SI_ERR_ARG, "si_fb_tune: jobs must share padded states and numeric space");
    }
    const size_t need = si_fb_workspace_bytes(jobs, n_jobs);
    if (need && (!workspace || workspace_bytes < need))
        return fail(SI_ERR_WORKSPACE, "si_fb_tune: workspace of %zu bytes needed, %zu given", need,
                    workspace_bytes);
    hipStream_t st = static_cast<hipStream_t>(stream_v);
    const uint32_t cap = max_markers ? max_markers : 2048u;
    const uint32_t Lp = si_padded_states(jobs[0].n_states);
    uint32_t gc[8], gb[8];
    const uint32_t nc = std::min(8u, si::fb_candidates(Lp, 0, gc, 8));
    const uint32_t nb = std::min(8u, si::fb_candidates(Lp, 1, gb, 8));
    if (!nc || !nb) return fail(SI_ERR_UNSUPPORTED, "si_fb_tune: no kernels for %u states", Lp);
    float chain_ms[kNumCkpt][8] = {}, block_ms[kNumCkpt][8] = {};
    Timer tm;
    if (!tm.ok) return device_status(SI_ERR_DEVICE, "si_fb_tune (events)");
    for (int k = 0; k < kNumCkpt; ++k) {
        const uint32_t cb = kCkptBlocks[k];
        for (uint32_t i = 0; i < n_jobs; ++i) {
            const si_fb_job& j = jobs[i];
            if (j.n_samples == 0 || j.n_markers == 0) continue;
            si::JobView v = tuning_view(j, cap, cb);
            si::fb_bind_workspace(v, Lp, static_cast<char*>(workspace));
            const si::FbTables t{j.eps, j.emax, j.palette, j.n_palette, si::fb_untyped_code(j)};
            for (uint32_t a = 0; a < nc; ++a) {
                const si::FbVariant var{gc[a], gb[0], cb};
                float ms = 0.0f;
                const int rc = tm.time(st, [&] { return si::launch_fb(v, t, Lp, SI_FB_PASS_CHAIN, var, st); }, &ms);
                if (rc) return device_status(rc, "si_fb_tune");
                chain_ms[k][a] += ms;
            }
            for (uint32_t a = 0; a < nb; ++a) {
                const si::FbVariant var{gc[0], gb[a], cb};
                float ms = 0.0f;
                const int rc = tm.time(st, [&] { return si::launch_fb(v, t, Lp, SI_FB_PASS_BLOCK, var, st); }, &ms);
                if (rc) return device_status(rc, "si_fb_tune");
                block_ms[k][a] += ms;
            }
        }
    }
    const si::FbVariant best = pick(chain_ms, gc, nc, block_ms, gb, nb, Lp);
    *out = si_fb_variant{best.gc, best.gb, best.cb};
    return SI_OK;
}

// ------------------------------------------------------------ batched form --
struct si_fb_batch {
    struct Class {
        uint32_t Lp = 0, flags = 0, pm = 0;
        std::vector<uint32_t> order;  // job indices, longest first
        // device tables (room for `cap` jobs; kept and reused by si_fb_batch_reset)
        si::JobView* d_jobs = nullptr;
        uint64_t* d_off = nullptr;    // chain, block and mask-pass offsets, [n+1] each
        si::JobView* d_tjobs = nullptr;  // tuning copies
        uint64_t* d_toff = nullptr;
        unsigned long long* d_counters = nullptr;
        size_t cap = 0;
        std::vector<uint64_t> h_off;  // offsets for d_off, uploaded by the next bind
        uint64_t chain_items = 0, block_items = 0, mask_groups = 0;
        uint64_t units = 0;           // sum of samples x markers over the class's jobs
        uint32_t off_cb = 0;          // checkpoint block the offsets were made for
        si::FbVariant var{};
        bool tuned = false, fixed = false;
        uint32_t grid_chain = 1, grid_block = 1;
    };
    std::vector<si_fb_job> jobs;
    std::vector<size_t> ws_off;  // per job
    size_t ws_bytes = 0;
    si::FbTables tables{};
    std::vector<Class> classes;
    // with several classes: a stream per class, forked from and joined back
    // into the caller's stream, so the classes' passes overlap
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev;  // [0] fork, [1 + k] class k done
    void* bound_ws = nullptr;
    hipGraphExec_t exec = nullptr;
    void* graph_ws = nullptr;
    // Page-locked staging of the device tables: a bind writes them here and
    // copies them on the run's stream; `staged` marks the copies' end, which
    // the next bind waits for before writing the staging again.
    char* h_stage = nullptr;
    size_t h_stage_bytes = 0;
    hipEvent_t staged = nullptr;
    // device tables of classes a reset dropped, reused by later classes
    std::vector<Class> spare;
};

static void class_mem_free(si_fb_batch::Class& c) {
    if (c.d_jobs) (void)hipFree(c.d_jobs);
    if (c.d_off) (void)hipFree(c.d_off);
    if (c.d_tjobs) (void)hipFree(c.d_tjobs);
    if (c.d_toff) (void)hipFree(c.d_toff);
    if (c.d_counters) (void)hipFree(c.d_counters);
    c.d_jobs = c.d_tjobs = nullptr;
    c.d_off = c.d_toff = nullptr;
    c.d_counters = nullptr;
    c.cap = 0;
}

static void batch_free(si_fb_batch* b) {
    if (!b) return;
    if (b->staged) (void)hipEventSynchronize(b->staged);
    for (auto& c : b->classes) class_mem_free(c);
    for (auto& c : b->spare) class_mem_free(c);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    if (b->staged) (void)hipEventDestroy(b->staged);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    for (hipStream_t s : b->side)
        if (s) (void)hipStreamDestroy(s);
    for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
    delete b;
}

extern "C++" {
// Item offsets of a class for checkpoint block cb (markers capped at cap).
static void class_offsets(const si_fb_batch& b, const si_fb_batch::Class& c, uint32_t cb,
                          uint32_t cap, std::vector<uint64_t>& off) {
    const size_t n = c.order.size();
    off.assign(3 * (n + 1), 0);
    for (size_t k = 0; k < n; ++k) {
        const si_fb_job& j = b.jobs[c.order[k]];
        const uint64_t M = std::min(j.n_markers, cap);
        const uint64_t nblk = (M + cb - 1) / cb;
        // samples padded to whole wavefront rounds (see fb_chain_batch_kernel,
        // fb_block_batch_kernel)
        const uint64_t Bp = si::fb_batch_padded_samples(j.n_samples);
        off[k + 1] = off[k] + 2ull * Bp;
        off[n + 1 + k + 1] = off[n + 1 + k] + Bp * nblk;
        off[2 * (n + 1) + k + 1] = off[2 * (n + 1) + k] + si::fb_tmask_groups(uint32_t(M));
    }
}

// Offsets and persistent grids of a class for its current variant.
static int class_prepare(si_fb_batch* b, si_fb_batch::Class& c) {
    if (c.off_cb != c.var.cb) {
        std::vector<uint64_t>& off = c.h_off;
        class_offsets(*b, c, c.var.cb, 0xffffffffu, off);
        const size_t n = c.order.size();
        c.chain_items = off[n];
        c.block_items = off[2 * n + 1];
        c.mask_groups = off[3 * n + 2];
        c.off_cb = c.var.cb;
        b->bound_ws = nullptr;  // new offsets, and job views carry the block count: rebind
        if (b->exec) {
            (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
        }
    }
    si::BatchClass bc{};
    bc.Lp = c.Lp;
    bc.flags = c.flags;
    bc.chain_items = c.chain_items;
    bc.block_items = c.block_items;
    si::fb_batch_grid(bc, c.var, b->tables.P);
    c.grid_chain = bc.grid_chain;
    c.grid_block = bc.grid_block;
    return SI_OK;
}
}  // extern "C++"

extern "C++" {
static int check_batch_jobs(const si_fb_job* jobs, uint32_t n_jobs, const char* what) {
    if (n_jobs == 0) return fail(SI_ERR_ARG, "%s: no jobs", what);
    for (uint32_t i = 0; i < n_jobs; ++i) {
        const int rc = check_job(jobs[i], i);
        if (rc) return rc;
        if (jobs[i].eps != jobs[0].eps || jobs[i].emax != jobs[0].emax ||
            jobs[i].palette != jobs[0].palette || jobs[i].n_palette != jobs[0].n_palette ||
            si::fb_untyped_code(jobs[i]) != si::fb_untyped_code(jobs[0]))
            return fail(SI_ERR_ARG, "%s: job %u uses a different emission table", what, i);
    }
    return SI_OK;
}

// Device tables for a class of n jobs: the smallest spare tables with room,
// else new ones (1.5x the need and at least 64 jobs, so a few resets of
// growing batches settle).  Tables too small stay spare until the batch is
// destroyed: freeing device memory can wait for every stream of the device,
// which a serving loop planning its next batch must not.
static int class_mem(si_fb_batch* b, si_fb_batch::Class& c, size_t n) {
    size_t best = b->spare.size();
    for (size_t k = 0; k < b->spare.size(); ++k)
        if (b->spare[k].cap >= n && (best == b->spare.size() || b->spare[k].cap < b->spare[best].cap)) best = k;
    if (best < b->spare.size()) {
        si_fb_batch::Class& sp = b->spare[best];
        c.d_jobs = sp.d_jobs;
        c.d_off = sp.d_off;
        c.d_tjobs = sp.d_tjobs;
        c.d_toff = sp.d_toff;
        c.d_counters = sp.d_counters;
        c.cap = sp.cap;
        b->spare.erase(b->spare.begin() + long(best));
        return SI_OK;
    }
    const size_t cap = std::max<size_t>({n + n / 2, n, 64});
    if (hipMalloc(&c.d_jobs, cap * sizeof(si::JobView)) != hipSuccess ||
        hipMalloc(&c.d_off, 3 * (cap + 1) * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&c.d_tjobs, cap * sizeof(si::JobView)) != hipSuccess ||
        hipMalloc(&c.d_toff, 3 * (cap + 1) * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&c.d_counters, 2 * sizeof(unsigned long long)) != hipSuccess) {
        class_mem_free(c);
        return SI_ERR_DEVICE;
    }
    c.cap = cap;
    return SI_OK;
}

// Plans `b` for `jobs`: per-job workspace offsets, classes by (padded states,
// numeric space, untyped-step kernels) with jobs longest chain first, device
// tables (the batch's own where they have room), side streams and events.
static int batch_plan(si_fb_batch* b, const si_fb_job* jobs, uint32_t n_jobs, const char* what) {
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    b->graph_ws = nullptr;
    b->bound_ws = nullptr;
    for (auto& c : b->classes) {
        si_fb_batch::Class sp;
        sp.d_jobs = c.d_jobs;
        sp.d_off = c.d_off;
        sp.d_tjobs = c.d_tjobs;
        sp.d_toff = c.d_toff;
        sp.d_counters = c.d_counters;
        sp.cap = c.cap;
        b->spare.push_back(std::move(sp));
    }
    b->classes.clear();
    b->jobs.assign(jobs, jobs + n_jobs);
    b->tables = si::FbTables{jobs[0].eps, jobs[0].emax, jobs[0].palette, jobs[0].n_palette,
                             si::fb_untyped_code(jobs[0])};
    b->ws_off.assign(n_jobs, 0);
    b->ws_bytes = 0;
    for (uint32_t i = 0; i < n_jobs; ++i) {
        b->ws_off[i] = b->ws_bytes;
        const size_t w = si::fb_workspace_bytes_one(jobs[i].n_samples, jobs[i].n_markers,
                                                    si_padded_states(jobs[i].n_states),
                                                    si::fb_window(jobs[i]), si::fb_band_job(jobs[i]));
        b->ws_bytes += (w + 255) & ~size_t(255);
    }
    for (uint32_t i = 0; i < n_jobs; ++i) {
        if (jobs[i].n_samples == 0 || jobs[i].n_markers == 0) continue;
        const uint32_t Lp = si_padded_states(jobs[i].n_states);
        const uint32_t fl = (jobs[i].flags & SI_FB_LOGSPACE) | (si::fb_band_job(jobs[i]) ? SI_FB_BANDED : 0u) |
                            (si::fb_fast_job(jobs[i]) ? si::kFbFastClass : 0u);
        auto it = std::find_if(b->classes.begin(), b->classes.end(),
                               [&](const si_fb_batch::Class& c) { return c.Lp == Lp && c.flags == fl; });
        if (it == b->classes.end()) {
            b->classes.emplace_back();
            it = b->classes.end() - 1;
            it->Lp = Lp;
            it->flags = fl;
            it->var = si::fb_default_variant(Lp);
        }
        it->order.push_back(i);
        if (jobs[i].flags & SI_FB_PER_MARKER) it->pm = 1;
    }
    for (auto& c : b->classes) {
        std::stable_sort(c.order.begin(), c.order.end(), [&](uint32_t x, uint32_t y) {
            return jobs[x].n_markers > jobs[y].n_markers;
        });
        for (uint32_t i : c.order) c.units += uint64_t(jobs[i].n_samples) * jobs[i].n_markers;
        if (class_mem(b, c, c.order.size()) != SI_OK || class_prepare(b, c) != SI_OK)
            return device_status(SI_ERR_DEVICE, what);
    }
    if (b->classes.size() > 1) {
        while (b->side.size() < b->classes.size()) {
            hipStream_t s = nullptr;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
                return device_status(SI_ERR_DEVICE, "batch side streams");
            b->side.push_back(s);
        }
        while (b->ev.size() < b->classes.size() + 1) {
            hipEvent_t e = nullptr;
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                return device_status(SI_ERR_DEVICE, "batch side events");
            b->ev.push_back(e);
        }
    }
    return SI_OK;
}
}  // extern "C++"

int si_fb_batch_create(const si_fb_job* jobs, uint32_t n_jobs, si_fb_batch** out) {
    if (!out || (n_jobs && !jobs)) return fail(SI_ERR_ARG, "si_fb_batch_create: bad arguments");
    *out = nullptr;
    int rc = check_batch_jobs(jobs, n_jobs, "si_fb_batch_create");
    if (rc) return rc;
    auto* b = new si_fb_batch();
    rc = batch_plan(b, jobs, n_jobs, "si_fb_batch_create");
    if (rc) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return SI_OK;
}

int si_fb_batch_reset(si_fb_batch* b, const si_fb_job* jobs, uint32_t n_jobs) {
    if (!b || (n_jobs && !jobs)) return fail(SI_ERR_ARG, "si_fb_batch_reset: bad arguments");
    const int rc = check_batch_jobs(jobs, n_jobs, "si_fb_batch_reset");
    if (rc) return rc;
    return batch_plan(b, jobs, n_jobs, "si_fb_batch_reset");
}

size_t si_fb_batch_workspace_bytes(const si_fb_batch* b) { return b ? b->ws_bytes : 0; }

uint32_t si_fb_batch_n_classes(const si_fb_batch* b) { return b ? uint32_t(b->classes.size()) : 0u; }

int si_fb_batch_class_info(const si_fb_batch* b, uint32_t k, uint32_t* padded_states,
                           uint32_t* flags, si_fb_variant* variant, int* tuned) {
    if (!b || k >= b->classes.size()) return fail(SI_ERR_ARG, "si_fb_batch_class_info: no class %u", k);
    const auto& c = b->classes[k];
    if (padded_states) *padded_states = c.Lp;
    if (flags) *flags = c.flags;
    if (variant) *variant = si_fb_variant{c.var.gc, c.var.gb, c.var.cb};
    if (tuned) *tuned = c.tuned ? 1 : 0;
    return SI_OK;
}

int si_fb_batch_set_variant(si_fb_batch* b, uint32_t k, const si_fb_variant* v) {
    if (!b || !v || k >= b->classes.size())
        return fail(SI_ERR_ARG, "si_fb_batch_set_variant: bad arguments");
    auto& c = b->classes[k];
    const int rc = check_variant(c.Lp, v, "si_fb_batch_set_variant");
    if (rc) return rc;
    const si::FbVariant nv{v->chain_lanes, v->block_lanes, v->ckpt_block};
    if (nv.gc != c.var.gc || nv.gb != c.var.gb || nv.cb != c.var.cb) {
        c.var = nv;
        if (b->exec) {
            (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
        }
    }
    c.fixed = true;
    c.tuned = false;
    if (class_prepare(b, c) != SI_OK) return device_status(SI_ERR_DEVICE, "si_fb_batch_set_variant");
    return SI_OK;
}

extern "C++" {
// Times every candidate of a class with the batched kernels on its jobs cut
// to `cap` markers, outputs disabled, checkpoints in the batch workspace.
static int class_tune(si_fb_batch* b, si_fb_batch::Class& c, char* ws, uint32_t cap,
                      hipStream_t st, bool side, si::FbVariant* best) {
    uint32_t gc[8], gb[8];
    const uint32_t nc = std::min(8u, si::fb_candidates(c.Lp, 0, gc, 8));
    const uint32_t nb = std::min(8u, si::fb_candidates(c.Lp, 1, gb, 8));
    if (!nc || !nb) return SI_ERR_UNSUPPORTED;
    float chain_ms[kNumCkpt][8] = {}, block_ms[kNumCkpt][8] = {};
    Timer tm;
    if (!tm.ok) return SI_ERR_DEVICE;
    const size_t n = c.order.size();
    // Job views cut to `mk` markers at checkpoint block cb, on the device.
    auto upload = [&](uint32_t cb, uint32_t mk, si::BatchClass& bc) -> int {
        std::vector<si::JobView> v(n);
        for (size_t i = 0; i < n; ++i) {
            const uint32_t ji = c.order[i];
            v[i] = tuning_view(b->jobs[ji], mk, cb);
            si::fb_bind_workspace(v[i], c.Lp, ws + b->ws_off[ji]);
        }
        std::vector<uint64_t> off;
        class_offsets(*b, c, cb, mk, off);
        if (hipMemcpy(c.d_tjobs, v.data(), n * sizeof(si::JobView), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c.d_toff, off.data(), off.size() * sizeof(uint64_t), hipMemcpyHostToDevice) != hipSuccess)
            return SI_ERR_DEVICE;
        bc = si::BatchClass{};
        bc.Lp = c.Lp;
        bc.n_jobs = uint32_t(n);
        bc.flags = c.flags;
        bc.jobs = c.d_tjobs;
        bc.chain_off = c.d_toff;
        bc.block_off = c.d_toff + n + 1;
        bc.mask_off = c.d_toff + 2 * (n + 1);
        bc.chain_items = off[n];
        bc.block_items = off[2 * n + 1];
        bc.mask_groups = off[3 * n + 2];
        bc.counters = c.d_counters;
        return 0;
    };
    for (int k = 0; k < kNumCkpt; ++k) {
        const uint32_t cb = kCkptBlocks[k];
        si::BatchClass bc{};
        // warm every candidate (code objects, caches) on 16-marker cuts
        int rc = upload(cb, 16, bc);
        if (rc) return rc;
        for (uint32_t a = 0; a < std::max(nc, nb) && !rc; ++a) {
            const si::FbVariant var{gc[std::min(a, nc - 1)], gb[std::min(a, nb - 1)], cb};
            si::fb_batch_grid(bc, var, b->tables.P);
            rc = si::launch_fb_batch(bc, b->tables, SI_FB_PASS_CHAIN | SI_FB_PASS_BLOCK, var, st);
        }
        if (rc || hipStreamSynchronize(st) != hipSuccess) return rc ? rc : SI_ERR_DEVICE;
        // then timed runs on the jobs cut to `cap`: every candidate twice, in
        // turn, keeping the faster (differences between candidates are a few
        // percent, about the spread of single runs)
        rc = upload(cb, cap, bc);
        if (rc) return rc;
        for (int round = 0; round < 2; ++round) {
            for (uint32_t a = 0; a < nc; ++a) {
                const si::FbVariant var{gc[a], gb[0], cb};
                si::fb_batch_grid(bc, var, b->tables.P);
                float ms = 0.0f;
                rc = tm.time_once(st, [&] { return si::launch_fb_batch(bc, b->tables, SI_FB_PASS_CHAIN, var, st); },
                                  &ms);
                if (rc) return rc;
                chain_ms[k][a] = round ? std::min(chain_ms[k][a], ms) : ms;
            }
            for (uint32_t a = 0; a < nb; ++a) {
                const si::FbVariant var{gc[0], gb[a], cb};
                si::fb_batch_grid(bc, var, b->tables.P);
                float ms = 0.0f;
                rc = tm.time_once(st, [&] { return si::launch_fb_batch(bc, b->tables, SI_FB_PASS_BLOCK, var, st); },
                                  &ms);
                if (rc) return rc;
                block_ms[k][a] = round ? std::min(block_ms[k][a], ms) : ms;
            }
        }
    }
    *best = pick(chain_ms, gc, nc, block_ms, gb, nb, c.Lp, side);
    return SI_OK;
}

}  // extern "C++"

void si_fb_tune_cache_clear(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache.clear();
}

int si_fb_batch_tune(si_fb_batch* b, void* workspace, size_t workspace_bytes, uint32_t max_markers,
                     void* stream_v) {
    if (!b) return fail(SI_ERR_ARG, "si_fb_batch_tune: null batch");
    if (b->ws_bytes && (!workspace || workspace_bytes < b->ws_bytes))
        return fail(SI_ERR_WORKSPACE, "si_fb_batch_tune: workspace of %zu bytes needed, %zu given",
                    b->ws_bytes, workspace_bytes);
    hipStream_t st = static_cast<hipStream_t>(stream_v);
    const uint32_t cap = max_markers ? max_markers : 0xffffffffu;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const bool side_by_side = b->classes.size() > 1 && !b->side.empty();
    auto key_of = [&](const si_fb_batch::Class& c) {
        return CacheKey{dev, c.Lp, c.flags, c.pm, size_bucket(c.chain_items), side_by_side ? 1u : 0u};
    };
    auto set_var = [&](si_fb_batch::Class& c, const si::FbVariant& v) {
        if (v.gc != c.var.gc || v.gb != c.var.gb || v.cb != c.var.cb) {
            c.var = v;
            if (b->exec) {
                (void)hipGraphExecDestroy(b->exec);
                b->exec = nullptr;
            }
        }
        return class_prepare(b, c);
    };
    // Several classes run side by side, so a class's chain pass is neither the
    // latency-bound pass it is alone nor a throughput-bound one: its lanes per
    // sample stay at the default (Lp / 4); the block pass and the checkpoint
    // block come from the class's own timing (see pick).
    for (auto& c : b->classes) {
        if (c.fixed) continue;
        si::FbVariant v{};
        if (!cache_get(key_of(c), &v)) {
            const int rc = class_tune(b, c, static_cast<char*>(workspace), cap, st, side_by_side, &v);
            if (rc) return device_status(rc, "si_fb_batch_tune");
        }
        c.tuned = true;
        if (set_var(c, v) != SI_OK) return device_status(SI_ERR_DEVICE, "si_fb_batch_tune");
    }
    for (const auto& c : b->classes)
        if (!c.fixed && c.tuned) cache_put(key_of(c), c.var);
    return SI_OK;
}

// Writes the device job tables (checkpoint pointers into `ws`) and item
// offsets: staged in page-locked memory and copied on `stream`, ahead of the
// launches that read them; the host waits for nothing but the previous
// bind's copies out of the same staging.
static int batch_bind(si_fb_batch* b, void* ws, hipStream_t stream) {
    size_t need = 0;
    for (const auto& c : b->classes)
        need += c.order.size() * sizeof(si::JobView) + c.h_off.size() * sizeof(uint64_t);
    if (b->staged && hipEventSynchronize(b->staged) != hipSuccess)
        return device_status(SI_ERR_DEVICE, "si_fb_batch_run (bind)");
    if (need > b->h_stage_bytes) {
        if (b->h_stage) (void)hipHostFree(b->h_stage);
        b->h_stage = nullptr;
        b->h_stage_bytes = 0;
        const size_t sz = std::max<size_t>(need + need / 2, 256u << 10);
        if (hipHostMalloc(reinterpret_cast<void**>(&b->h_stage), sz, 0) != hipSuccess)
            return device_status(SI_ERR_DEVICE, "si_fb_batch_run (bind staging)");
        b->h_stage_bytes = sz;
    }
    if (!b->staged && hipEventCreateWithFlags(&b->staged, hipEventDisableTiming) != hipSuccess) {
        b->staged = nullptr;
        return device_status(SI_ERR_DEVICE, "si_fb_batch_run (bind event)");
    }
    char* p = b->h_stage;
    for (auto& c : b->classes) {
        auto* v = reinterpret_cast<si::JobView*>(p);
        for (size_t k = 0; k < c.order.size(); ++k) {
            const uint32_t i = c.order[k];
            v[k] = si::fb_job_view(b->jobs[i], c.var.cb);
            si::fb_bind_workspace(v[k], c.Lp, static_cast<char*>(ws) + b->ws_off[i]);
        }
        const size_t jb = c.order.size() * sizeof(si::JobView), ob = c.h_off.size() * sizeof(uint64_t);
        std::memcpy(p + jb, c.h_off.data(), ob);
        if (hipMemcpyAsync(c.d_jobs, p, jb, hipMemcpyHostToDevice, stream) != hipSuccess ||
            hipMemcpyAsync(c.d_off, p + jb, ob, hipMemcpyHostToDevice, stream) != hipSuccess)
            return device_status(SI_ERR_DEVICE, "si_fb_batch_run (bind)");
        p += jb + ob;
    }
    if (hipEventRecord(b->staged, stream) != hipSuccess)
        return device_status(SI_ERR_DEVICE, "si_fb_batch_run (bind)");
    b->bound_ws = ws;
    return SI_OK;
}

// Every class's chain pass lasts about as long as its longest chains take one
// wavefront, far below what its work needs of the GPU, so the classes run side
// by side on their own streams, the persistent chain grids shared out by the
// classes' chain work (lanes x samples x markers); each class's block pass
// follows its chain pass on its stream with the full grid.  `bcs` holds each
// class's queues (job tables, item counts, full persistent grids for its
// variant in `vars`).
static int enqueue_classes(si_fb_batch* b, std::vector<si::BatchClass> bcs,
                           const std::vector<si::FbVariant>& vars, hipStream_t stream) {
    const bool fork = b->classes.size() > 1 && !b->side.empty();
    double work_total = 0.0;
    for (size_t k = 0; k < b->classes.size(); ++k) work_total += double(b->classes[k].units) * vars[k].gc;
    if (fork && hipEventRecord(b->ev[0], stream) != hipSuccess) return SI_ERR_DEVICE;
    for (size_t k = 0; k < b->classes.size(); ++k) {
        si::BatchClass& bc = bcs[k];
        hipStream_t st = stream;
        if (fork) {
            st = b->side[k];
            if (hipStreamWaitEvent(st, b->ev[0], 0) != hipSuccess) return SI_ERR_DEVICE;
            const double share =
                work_total > 0.0 ? double(b->classes[k].units) * vars[k].gc / work_total : 1.0;
            bc.grid_chain = std::max<uint32_t>(1u, std::min<uint32_t>(
                bc.grid_chain, uint32_t(std::ceil(double(bc.grid_chain) * share))));
        }
        const int rc = si::launch_fb_batch(bc, b->tables, SI_FB_PASS_CHAIN | SI_FB_PASS_BLOCK, vars[k], st);
        if (rc) return rc;
        if (fork && hipEventRecord(b->ev[1 + k], st) != hipSuccess) return SI_ERR_DEVICE;
    }
    if (fork)
        for (size_t k = 0; k < b->classes.size(); ++k)
            if (hipStreamWaitEvent(stream, b->ev[1 + k], 0) != hipSuccess) return SI_ERR_DEVICE;
    return SI_OK;
}

static int batch_enqueue(si_fb_batch* b, hipStream_t stream) {
    std::vector<si::BatchClass> bcs(b->classes.size());
    std::vector<si::FbVariant> vars(b->classes.size());
    for (size_t k = 0; k < b->classes.size(); ++k) {
        const auto& c = b->classes[k];
        si::BatchClass& bc = bcs[k];
        bc.Lp = c.Lp;
        bc.n_jobs = uint32_t(c.order.size());
        bc.flags = c.flags;
        bc.jobs = c.d_jobs;
        bc.chain_off = c.d_off;
        bc.block_off = c.d_off + c.order.size() + 1;
        bc.mask_off = c.d_off + 2 * (c.order.size() + 1);
        bc.chain_items = c.chain_items;
        bc.block_items = c.block_items;
        bc.mask_groups = c.mask_groups;
        bc.counters = c.d_counters;
        bc.grid_chain = c.grid_chain;
        bc.grid_block = c.grid_block;
        vars[k] = c.var;
    }
    const int rc = enqueue_classes(b, bcs, vars, stream);
    return rc ? device_status(rc, "si_fb_batch_run") : SI_OK;
}

int si_fb_batch_run(si_fb_batch* b, void* workspace, size_t workspace_bytes, int use_graph,
                    void* stream_v) {
    if (!b) return fail(SI_ERR_ARG, "si_fb_batch_run: null batch");
    if (b->ws_bytes && (!workspace || workspace_bytes < b->ws_bytes))
        return fail(SI_ERR_WORKSPACE, "si_fb_batch_run: workspace of %zu bytes needed, %zu given",
                    b->ws_bytes, workspace_bytes);
    hipStream_t stream = static_cast<hipStream_t>(stream_v);
    for (const auto& c : b->classes)
        if (!c.fixed && !c.tuned) {
            const int rc = si_fb_batch_tune(b, workspace, workspace_bytes, 0, stream_v);
            if (rc) return rc;
            break;
        }
    if (workspace != b->bound_ws) {
        const int rc = batch_bind(b, workspace, stream);
        if (rc) return rc;
    }
    if (!use_graph) return batch_enqueue(b, stream);
    if (!stream) return fail(SI_ERR_ARG, "si_fb_batch_run: graph capture needs a non-null stream");
    if (!b->exec || b->graph_ws != workspace) {
        if (b->exec) {
            (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
        }
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return device_status(SI_ERR_DEVICE, "si_fb_batch_run (capture)");
        const int rc = batch_enqueue(b, stream);
        const hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess || hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return device_status(SI_ERR_DEVICE, "si_fb_batch_run (instantiate)");
        }
        (void)hipGraphDestroy(g);
        b->graph_ws = workspace;
    }
    if (hipGraphLaunch(b->exec, stream) != hipSuccess)
        return device_status(SI_ERR_DEVICE, "si_fb_batch_run (graph launch)");
    return SI_OK;
}

int si_fb_batch_destroy(si_fb_batch* b) {
    batch_free(b);
    return SI_OK;
}

float si_sample_uniform(uint64_t seed, uint32_t s, uint32_t m, uint32_t rng_stream) {
    uint32_t c0 = m >> 2, c1 = s, c2 = rng_stream, c3 = 0;
    uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0, p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0, n1 = uint32_t(p1);
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1, n3 = uint32_t(p0);
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    const uint32_t w[4] = {c0, c1, c2, c3};
    return float(w[m & 3u] >> 8) * (1.0f / 16777216.0f);
}

}  // extern "C"
