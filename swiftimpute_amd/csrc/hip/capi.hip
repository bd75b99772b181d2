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
#include <cstring>
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
    (void)H;
    (void)M;
    return 0;
}

int si_pbwt_sweep(const uint64_t* cols, uint32_t H, uint32_t M, uint32_t* prefix, uint32_t* div,
                  void* workspace, size_t workspace_bytes, void* stream) {
    (void)workspace;
    (void)workspace_bytes;
    if (H == 0) return fail(SI_ERR_ARG, "si_pbwt_sweep: H must be positive");
    if (M == 0) return SI_OK;
    if (!cols || !prefix) return fail(SI_ERR_ARG, "si_pbwt_sweep: cols and prefix are required");
    return device_status(
        si::launch_pbwt_sweep(cols, H, M, prefix, div, static_cast<hipStream_t>(stream)),
        "si_pbwt_sweep");
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

int si_state_table(const uint32_t* prefix, const uint64_t* cols, const uint64_t* miss, uint32_t H,
                   uint32_t M, uint32_t L, uint32_t mode, const uint32_t* perm_host,
                   uint32_t* states, uint8_t* sa, uint8_t* mmask, void* stream) {
    if (H == 0 || L == 0 || L > 64)
        return fail(SI_ERR_ARG, "si_state_table: need H > 0 and 1 <= L <= 64 (got H=%u L=%u)", H, L);
    if (mode > 1) return fail(SI_ERR_ARG, "si_state_table: unknown mode %u", mode);
    if (mode == 1 && 2ull * L > H)
        return fail(SI_ERR_ARG, "si_state_table: simple selection needs 2L <= H");
    if (mode == 0 && (!prefix || !perm_host))
        return fail(SI_ERR_ARG, "si_state_table: PBWT selection needs prefix and perm");
    if (!cols || !sa || !mmask) return fail(SI_ERR_ARG, "si_state_table: cols, sa, mmask required");
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
    if (j.n_states == 0 || j.n_states > 64)
        return fail(SI_ERR_ARG, "job %u: n_states must be in 1..64 (got %u)", idx, j.n_states);
    if (j.n_palette == 0 || j.n_palette > si::kMaxPalette)
        return fail(SI_ERR_UNSUPPORTED, "job %u: palette size %u outside 1..%u", idx, j.n_palette,
                    si::kMaxPalette);
    if (j.code_pitch < j.n_markers || (j.code_pitch & 15u))
        return fail(SI_ERR_ARG, "job %u: code_pitch must be >= M and a multiple of 16", idx);
    if (j.n_samples == 0 || j.n_markers == 0) return SI_OK;
    const bool logspace = (j.flags & SI_FB_LOGSPACE) != 0;
    if (!j.sa || !j.mmask || !j.codes)
        return fail(SI_ERR_ARG, "job %u: sa, mmask, codes are required", idx);
    if (logspace ? !j.palette : (!j.eps || !j.emax))
        return fail(SI_ERR_ARG, "job %u: %s required", idx,
                    logspace ? "palette is" : "eps and emax are");
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
    return SI_OK;
}

size_t si_fb_workspace_bytes(const si_fb_job* jobs, uint32_t n_jobs) {
    size_t mx = 0;
    for (uint32_t i = 0; jobs && i < n_jobs; ++i)
        mx = std::max(mx, si::fb_workspace_bytes_one(jobs[i].n_samples, jobs[i].n_markers,
                                                     si_padded_states(jobs[i].n_states)));
    return mx;
}

int si_fb_run(const si_fb_job* jobs, uint32_t n_jobs, void* workspace, size_t workspace_bytes,
              void* stream) {
    return si_fb_run_passes(jobs, n_jobs, workspace, workspace_bytes,
                            SI_FB_PASS_CHAIN | SI_FB_PASS_BLOCK, stream);
}

int si_fb_run_passes(const si_fb_job* jobs, uint32_t n_jobs, void* workspace,
                     size_t workspace_bytes, uint32_t passes, void* stream) {
    if (n_jobs && !jobs) return fail(SI_ERR_ARG, "si_fb_run: jobs is null");
    if (passes == 0 || (passes & ~3u))
        return fail(SI_ERR_ARG, "si_fb_run_passes: passes must be a nonzero CHAIN|BLOCK mask");
    // Jobs share one workspace in stream order, so a split sequence
    // (all chains, then all blocks) is only meaningful for a single job.
    if (passes != 3u && n_jobs > 1)
        return fail(SI_ERR_ARG, "si_fb_run_passes: a single pass runs one job at a time");
    for (uint32_t i = 0; i < n_jobs; ++i) {
        const int rc = check_job(jobs[i], i);
        if (rc) return rc;
    }
    const size_t need = si_fb_workspace_bytes(jobs, n_jobs);
    if (need && (!workspace || workspace_bytes < need))
        return fail(SI_ERR_WORKSPACE, "si_fb_run: workspace of %zu bytes needed, %zu given", need,
                    workspace_bytes);
    for (uint32_t i = 0; i < n_jobs; ++i) {
        const si_fb_job& j = jobs[i];
        if (j.n_samples == 0 || j.n_markers == 0) continue;
        si::FbParams p{};
        p.B = j.n_samples;
        p.M = j.n_markers;
        p.L = j.n_states;
        p.Lp = si_padded_states(j.n_states);
        p.flags = j.flags;
        p.pitch = j.code_pitch;
        p.P = j.n_palette;
        p.rng_stream = j.rng_stream;
        p.sample_base = j.sample_base;
        p.nblk = (j.n_markers + si::kBlock - 1) / si::kBlock;
        p.seed = j.seed;
        p.step = j.step;
        p.steps = j.steps;
        p.sa = j.sa;
        p.mmask = j.mmask;
        p.codes = j.codes;
        p.eps = j.eps;
        p.emax = j.emax;
        p.palette = j.palette;
        p.ds = j.ds;
        p.ap = j.ap;
        p.hap = j.hap;
        p.prob = j.prob;
        p.loglik = j.loglik;
        char* ws = static_cast<char*>(workspace);
        const size_t vec = size_t(p.B) * p.nblk * p.Lp * sizeof(float);
        const size_t vec_al = (vec + 255) & ~size_t(255);
        p.ck_alpha = reinterpret_cast<float*>(ws);
        p.ck_beta = reinterpret_cast<float*>(ws + vec_al);
        p.ck_bl = reinterpret_cast<double*>(ws + 2 * vec_al);
        const int rc = si::launch_fb(p, passes, static_cast<hipStream_t>(stream));
        if (rc) return device_status(rc, "si_fb_run");
    }
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
