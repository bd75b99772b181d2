// si_oracle.cpp -- CPU restatement of SwiftImpute's imputation hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (swiftimpute_amd/, the HIP
// library, the C++ API) links, loads or calls this file.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
// checker and as the timed CPU baseline.
//
// Every function names the reference file:line it restates (paths relative to
// the tcBio/SwiftImpute tree).  Pinning status of each function (what checks it
// independently of this restatement) is recorded in the header comment of the
// function and exercised by tests/test_oracle_*.py:
//
//   or_pbwt_build          pinned: hand KAT on test_reference.vcf, brute-force
//                          lexicographic sort of reversed prefixes, invariants.
//   or_select_perm         pinned: libstdc++ std::sort is the library routine the
//                          reference calls; KATs of the permutation prefixes.
//   or_select_states_sort  pinned: equals or_select_perm gather (div == m).
//   or_select_states_gpu_* pinned: closed forms with div == m ([p0, p32, p64,
//                          p96, 4, 5, ...] and p[0..L-1]), hand-traced rounds
//                          on crafted divergences, a stable sort by divergence
//                          for the insertion kernel on random rows.
//   or_transitions         pinned: closed forms (rho = 1 limit, d = 0 limit).
//   or_fb_f64 / _fast      pinned: brute-force path enumeration (M <= 6, L <= 3)
//                          with the -999 floor folded in, uniform-T closed form,
//                          the reference's own FB integration case.
//   or_fb_f32_faithful     pinned against or_fb_f64 (drift bound only).
//   or_fb_f64_windowed     pinned: equals or_fb_f64 when M <= W; with windows,
//                          brute-force path enumeration of the model it defines
//                          (tiny M), and window-local closed forms (a window
//                          whose forward restarts at a fixed vector).
//   or_philox4x32_10       pinned: Random123 published KAT vectors and rocRAND's
//                          open-source header.
//   or_sample_from_probs   pinned: closed form on hand-built vectors.
//
// Build: see oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <utility>
#include <vector>

namespace {

constexpr double kLogZero = -999.0;   // forward_backward.cu:22 (LOG_ZERO)
constexpr float kLogZeroF = -999.0f;

// ---------------------------------------------------------------------------
// Log-sum-exp of two values, natural log.  Restates logsumexp2
// (src/kernels/logsumexp.cuh:67-70) in the given precision.
// ---------------------------------------------------------------------------
inline double lse2(double a, double b) {
    double mx = std::fmax(a, b);
    return std::log(std::exp(a - mx) + std::exp(b - mx)) + mx;
}
inline float lse2f(float a, float b) {
    float mx = std::fmax(a, b);
    return std::log(std::exp(a - mx) + std::exp(b - mx)) + mx;
}

// Emission lookup: compute_single_emission (src/kernels/emission.cu:35-54).
// The allele sum is an int, so a missing allele (255) lands in the ll_11 branch.
inline float emission_lookup(const float* gl3, uint8_t a0, uint8_t a1) {
    int sum = int(a0) + int(a1);
    if (sum == 0) return gl3[0];
    if (sum == 1) return gl3[1];
    return gl3[2];
}

// Index of the transition interval used for the step INTO marker m (m >= 1).
// Compatibility mode reproduces the reference, whose kernels index
// transition_matrix[from*L+to] with no marker offset
// (forward_backward.cu:108,201,234,269): interval 0 for every step.
inline uint32_t interval_for_step(uint32_t m, int per_marker) {
    return per_marker ? (m - 1) : 0u;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// A1: PBWT sweep.  Restates PBWTBuilder::build_parallel + build_marker
// (src/pbwt/pbwt_index.cpp:153-208, 210-247): prev = identity; for each marker
// the previous order is split stably into allele==0 and allele!=0 lists; the
// divergence pushed is p/q which are reset to m after every push, so every
// divergence in row m equals m (pbwt_index.cpp:173-192).
// ref: [M][H] uint8, marker-major (imputer.hpp:96).  Outputs [M][H] uint32.
// ---------------------------------------------------------------------------
int or_pbwt_build(const uint8_t* ref, uint32_t M, uint32_t H,
                  uint32_t* prefix, uint32_t* div) {
    if (H == 0) return -1;
    std::vector<uint32_t> prev(H), prev_div(H, 0);
    std::iota(prev.begin(), prev.end(), 0u);
    std::vector<uint32_t> z, zd, o, od;
    z.reserve(H); zd.reserve(H); o.reserve(H); od.reserve(H);
    for (uint32_t m = 0; m < M; ++m) {
        z.clear(); zd.clear(); o.clear(); od.clear();
        uint32_t p = m, q = m;
        const uint8_t* col = ref + size_t(m) * H;
        for (uint32_t i = 0; i < H; ++i) {
            uint32_t hap = prev[i];
            if (col[hap] == 0) { z.push_back(hap); zd.push_back(p); p = m; }
            else               { o.push_back(hap); od.push_back(q); q = m; }
        }
        uint32_t* row = prefix + size_t(m) * H;
        uint32_t* drow = div ? div + size_t(m) * H : nullptr;
        size_t idx = 0;
        for (size_t i = 0; i < z.size(); ++i, ++idx) { row[idx] = z[i]; if (drow) drow[idx] = zd[i]; }
        for (size_t i = 0; i < o.size(); ++i, ++idx) { row[idx] = o[i]; if (drow) drow[idx] = od[i]; }
        std::memcpy(prev.data(), row, sizeof(uint32_t) * H);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// A1, streamed: the same sweep over markers [m_base, m_base + M) of a panel,
// continuing from `order` (the prefix row after marker m_base - 1, or the
// identity when m_base == 0), which is updated to the row after the window's
// last marker.  Divergence entries are the global marker index
// (pbwt_index.cpp:173-192 reset p/q to m after every push).  ref holds the
// window's M rows [M][H]; prefix/div receive the window's rows (div may be
// null).  Splitting a panel into windows gives or_pbwt_build's rows exactly.
// ---------------------------------------------------------------------------
int or_pbwt_sweep_window(const uint8_t* ref, uint32_t M, uint32_t H, uint32_t m_base,
                         uint32_t* order, uint32_t* prefix, uint32_t* div) {
    if (H == 0 || !order) return -1;
    std::vector<uint32_t> z, o;
    z.reserve(H); o.reserve(H);
    for (uint32_t m = 0; m < M; ++m) {
        z.clear(); o.clear();
        const uint8_t* col = ref + size_t(m) * H;
        for (uint32_t i = 0; i < H; ++i) {
            const uint32_t hap = order[i];
            if (col[hap] == 0) z.push_back(hap);
            else o.push_back(hap);
        }
        std::memcpy(order, z.data(), sizeof(uint32_t) * z.size());
        std::memcpy(order + z.size(), o.data(), sizeof(uint32_t) * o.size());
        std::memcpy(prefix + size_t(m) * H, order, sizeof(uint32_t) * H);
        if (div) std::fill(div + size_t(m) * H, div + size_t(m + 1) * H, m_base + m);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// A2: the permutation libstdc++'s std::sort applies to H candidates with equal
// keys under the reference comparator (a.first > b.first,
// src/pbwt/pbwt_index.cpp:30-33).  perm[k] = original position of the element
// that ends at position k, for k < min(L, H).
// ---------------------------------------------------------------------------
int or_select_perm(uint32_t H, uint32_t L, uint32_t* perm) {
    if (H == 0) return -1;
    std::vector<std::pair<uint32_t, uint32_t>> cand(H);
    for (uint32_t i = 0; i < H; ++i) cand[i] = {0u, i};
    std::sort(cand.begin(), cand.end(),
              [](const std::pair<uint32_t, uint32_t>& a,
                 const std::pair<uint32_t, uint32_t>& b) { return a.first > b.first; });
    uint32_t n = std::min(L, H);
    for (uint32_t k = 0; k < n; ++k) perm[k] = cand[k].second;
    return 0;
}

// ---------------------------------------------------------------------------
// A2 literal: PBWTIndex::select_states + select_states_at_marker
// (src/pbwt/pbwt_index.cpp:20-44, 58-79) for one marker: candidates
// (div, hap) in prefix order, std::sort descending by div, first min(L,H),
// then fill i % H for i in [H, L).
// ---------------------------------------------------------------------------
int or_select_states_sort(const uint32_t* prefix_row, const uint32_t* div_row,
                          uint32_t H, uint32_t L, uint32_t* out) {
    if (H == 0) return -1;
    std::vector<std::pair<uint32_t, uint32_t>> cand;
    cand.reserve(H);
    for (uint32_t i = 0; i < H; ++i) cand.push_back({div_row[i], prefix_row[i]});
    std::sort(cand.begin(), cand.end(),
              [](const std::pair<uint32_t, uint32_t>& a,
                 const std::pair<uint32_t, uint32_t>& b) { return a.first > b.first; });
    for (uint32_t i = 0; i < L && i < cand.size(); ++i) out[i] = cand[i].second;
    for (uint32_t i = uint32_t(cand.size()); i < L; ++i) out[i] = i % H;
    return 0;
}

// ---------------------------------------------------------------------------
// A3: select_states_warp_kernel (src/pbwt/pbwt_selector.cu:212-312) for one
// (sample, marker), lane by lane.  Each of 32 lanes walks prefix positions
// lane, lane+32, ... keeping a top-4 by divergence with the kernel's
// insertion rule (:243-259); then round r = 0.. while r < L and r < the
// lane's count (:269) takes the warp maximum of every lane's r-th entry, the
// lowest lane holding it wins (ballot + ffs, :279-283), lane 0 writes the
// winner's haplotype and the winner zeroes its r-th divergence (:292-294).
// Slots from lane 0's count up to L are filled with i % H (:305-310).  The
// loop bound is per lane, so for H < 128 lanes leave before lane 0 and the
// kernel's full-mask shuffles are undefined; here only lanes still in the
// loop take part (lane 0 always is, holding the most entries).
// ---------------------------------------------------------------------------
int or_select_states_gpu_warp(const uint32_t* prefix_row, const uint32_t* div_row,
                              uint32_t H, uint32_t L, uint32_t* out) {
    if (H == 0 || L == 0) return -1;
    uint32_t bd[32][4] = {}, bh[32][4] = {}, cnt[32] = {};
    for (uint32_t lane = 0; lane < 32; ++lane) {
        for (uint32_t i = lane; i < H; i += 32) {
            const uint32_t hap = prefix_row[i], dv = div_row[i];
            for (uint32_t j = 0; j < 4; ++j) {
                if (cnt[lane] <= j || dv > bd[lane][j]) {
                    for (uint32_t k = 3; k > j; --k) {
                        bd[lane][k] = bd[lane][k - 1];
                        bh[lane][k] = bh[lane][k - 1];
                    }
                    bd[lane][j] = dv;
                    bh[lane][j] = hap;
                    cnt[lane] = std::min(cnt[lane] + 1, 4u);
                    break;
                }
            }
        }
    }
    for (uint32_t r = 0; r < L && r < cnt[0]; ++r) {
        uint32_t mx = 0;
        for (uint32_t lane = 0; lane < 32; ++lane)
            if (r < cnt[lane]) mx = std::max(mx, bd[lane][r]);
        uint32_t win = 0;
        while (!(r < cnt[win] && bd[win][r] == mx)) ++win;
        out[r] = bh[win][r];
        bd[win][r] = 0;
    }
    for (uint32_t i = cnt[0]; i < L; ++i) out[i] = i % H;
    return 0;
}

// ---------------------------------------------------------------------------
// A3: select_states_bitonic_kernel (src/pbwt/pbwt_selector.cu:320-394) for
// one marker: a linear scan keeping the top L by divergence, inserting before
// the first strictly smaller entry (:349-379); slots past the count are
// filled with i % H (:386-392).  L <= 64 (:340).
// ---------------------------------------------------------------------------
int or_select_states_gpu_insert(const uint32_t* prefix_row, const uint32_t* div_row,
                                uint32_t H, uint32_t L, uint32_t* out) {
    if (H == 0 || L == 0 || L > 64) return -1;
    uint32_t td[64], th[64], n = 0;
    for (uint32_t i = 0; i < H; ++i) {
        const uint32_t hap = prefix_row[i], dv = div_row[i];
        uint32_t pos;
        if (n < L) {
            pos = n;
            for (uint32_t j = 0; j < n; ++j)
                if (dv > td[j]) { pos = j; break; }
            for (uint32_t j = n; j > pos; --j) { td[j] = td[j - 1]; th[j] = th[j - 1]; }
            ++n;
        } else if (dv > td[L - 1]) {
            pos = L - 1;
            for (uint32_t j = 0; j + 1 < L; ++j)
                if (dv > td[j]) { pos = j; break; }
            for (uint32_t j = L - 1; j > pos; --j) { td[j] = td[j - 1]; th[j] = th[j - 1]; }
        } else {
            continue;
        }
        td[pos] = dv;
        th[pos] = hap;
    }
    for (uint32_t i = 0; i < L; ++i) out[i] = i < n ? th[i] : i % H;
    return 0;
}

// ---------------------------------------------------------------------------
// A3: select_states_kernel (src/pbwt/pbwt_selector.cu:47-204) for one marker,
// L > 64.  256 threads; thread t scans positions [t*c, min(t*c + c, H)),
// c = ceil(H/256): its first L entries go into a local list with the same
// insertion (before the first strictly smaller divergence, :84-105), later
// ones only when strictly better than the list's last (:108-128).  Thread 0
// then merges the lists in thread order with the same rule (:139-186) and
// fills slots past the count with i % H (:189-199).  A thread whose range
// holds fewer than L entries leaves the rest of its slots unwritten and the
// merge reads them (undefined in the reference); here they hold nothing.
// The reference keeps its merge list in a MAX_STATES_PER_BLOCK = 256 array.
// ---------------------------------------------------------------------------
namespace {
void top_insert(std::vector<uint32_t>& td, std::vector<uint32_t>& th, uint32_t L, uint32_t dv,
                uint32_t hap) {
    const uint32_t n = uint32_t(td.size());
    if (n < L) {
        uint32_t pos = n;
        for (uint32_t j = 0; j < n; ++j)
            if (dv > td[j]) { pos = j; break; }
        td.insert(td.begin() + pos, dv);
        th.insert(th.begin() + pos, hap);
    } else if (dv > td[L - 1]) {
        uint32_t pos = L - 1;
        for (uint32_t j = 0; j + 1 < L; ++j)
            if (dv > td[j]) { pos = j; break; }
        td.insert(td.begin() + pos, dv);
        th.insert(th.begin() + pos, hap);
        td.pop_back();
        th.pop_back();
    }
}
}  // namespace

int or_select_states_gpu_block(const uint32_t* prefix_row, const uint32_t* div_row, uint32_t H,
                               uint32_t L, uint32_t* out) {
    if (H == 0 || L == 0 || L > 256) return -1;
    constexpr uint32_t kThreads = 256;
    const uint32_t c = (H + kThreads - 1) / kThreads;
    std::vector<uint32_t> gd, gh;
    for (uint32_t t = 0; t < kThreads; ++t) {
        std::vector<uint32_t> td, th;
        const uint32_t b = t * c, e = std::min(H, b + c);
        for (uint32_t i = b; i < e; ++i) top_insert(td, th, L, div_row[i], prefix_row[i]);
        for (size_t k = 0; k < td.size(); ++k) top_insert(gd, gh, L, td[k], th[k]);
    }
    for (uint32_t i = 0; i < L; ++i) out[i] = i < gh.size() ? gh[i] : i % H;
    return 0;
}

// ---------------------------------------------------------------------------
// A3 dispatch: launch_select_states (src/pbwt/pbwt_selector.cu:400-463) picks
// the warp kernel for L <= 32 and H <= 10000, the insertion kernel for L <= 64
// and the block kernel above.
// ---------------------------------------------------------------------------
int or_select_states_gpu(const uint32_t* prefix_row, const uint32_t* div_row, uint32_t H,
                         uint32_t L, uint32_t* out) {
    if (L <= 32 && H <= 10000) return or_select_states_gpu_warp(prefix_row, div_row, H, L, out);
    if (L <= 64) return or_select_states_gpu_insert(prefix_row, div_row, H, L, out);
    return or_select_states_gpu_block(prefix_row, div_row, H, L, out);
}

// ---------------------------------------------------------------------------
// State table from the reference's GPU selection on the given PBWT arrays,
// expanded into pairs (h, (h+1) % H) as imputer.cpp:862-873 does, with the
// alleles looked up as in or_state_table.
// ---------------------------------------------------------------------------
int or_state_table_gpu(const uint8_t* ref, const uint32_t* prefix, const uint32_t* div,
                       uint32_t M, uint32_t H, uint32_t L, uint32_t* h0_out, uint32_t* h1_out,
                       uint8_t* a0_out, uint8_t* a1_out) {
    if (H == 0 || L == 0 || L > 256) return -1;
    std::vector<uint32_t> sel(L);
    for (uint32_t m = 0; m < M; ++m) {
        const int rc = or_select_states_gpu(prefix + size_t(m) * H, div + size_t(m) * H, H, L,
                                            sel.data());
        if (rc) return rc;
        for (uint32_t k = 0; k < L; ++k) {
            const uint32_t h0 = sel[k], h1 = (h0 + 1) % H;
            const size_t o = size_t(m) * L + k;
            if (h0_out) h0_out[o] = h0;
            if (h1_out) h1_out[o] = h1;
            a0_out[o] = ref[size_t(m) * H + h0];
            a1_out[o] = ref[size_t(m) * H + h1];
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// A2 + A4 + allele lookup: the per-marker state table.  states[m][k] =
// prefix[m][perm[k]] for k < min(L,H), else k % H (pbwt_index.cpp:36-43);
// partner = (h+1) % H (imputer.cpp:918-919).  simple_mode restates the
// use_pbwt_selection=false branch (imputer.cpp:933-942): (2k, 2k+1).
// Outputs: h0/h1 [M][L] uint32 (may be null) and raw alleles a0/a1 [M][L].
// ---------------------------------------------------------------------------
int or_state_table(const uint8_t* ref, const uint32_t* prefix, uint32_t M, uint32_t H,
                   uint32_t L, int simple_mode, uint32_t* h0_out, uint32_t* h1_out,
                   uint8_t* a0_out, uint8_t* a1_out) {
    if (H == 0 || L == 0) return -1;
    std::vector<uint32_t> perm(std::min(L, H));
    if (!simple_mode) or_select_perm(H, L, perm.data());
    else if (2ull * L > H) return -2;   // (2k+1) would index past the panel
    for (uint32_t m = 0; m < M; ++m) {
        for (uint32_t k = 0; k < L; ++k) {
            uint32_t h0, h1;
            if (simple_mode) { h0 = 2 * k; h1 = 2 * k + 1; }
            else {
                h0 = (k < H) ? prefix[size_t(m) * H + perm[k]] : (k % H);
                h1 = (h0 + 1) % H;
            }
            size_t o = size_t(m) * L + k;
            if (h0_out) h0_out[o] = h0;
            if (h1_out) h1_out[o] = h1;
            a0_out[o] = ref[size_t(m) * H + h0];
            a1_out[o] = ref[size_t(m) * H + h1];
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// A5: genetic map.  g[m] = cM/100 if cM > 0 else pos/1e8 Morgans
// (src/api/imputer.cpp:596-605).  cM may be null (all zero, as load_vcf sets).
// ---------------------------------------------------------------------------
void or_genetic_map(const uint64_t* pos, const double* cM, uint32_t M, double* g) {
    for (uint32_t m = 0; m < M; ++m) {
        double c = cM ? cM[m] : 0.0;
        g[m] = (c > 0.0) ? c / 100.0 : static_cast<double>(pos[m]) / 1e8;
    }
}

// ---------------------------------------------------------------------------
// A5: Li-Stephens transition constants per interval t = 0..M-2
// (src/kernels/transition.cu:83-121): FP64 math, stored FP32.  Td = diagonal
// (stay), To = off-diagonal (switch).  ne/rho_rate arrive as the reference
// copies them (uint32 ne and float rho_rate widened to double,
// imputer.cpp:608-612, transition.hpp:12-19).
// ---------------------------------------------------------------------------
void or_transitions(const double* g, uint32_t M, uint32_t ne_u, float rho_rate_f,
                    float* Td, float* To) {
    const double ne = static_cast<double>(ne_u);
    const double rho_rate = static_cast<double>(rho_rate_f);
    for (uint32_t t = 0; t + 1 < M; ++t) {
        double d = g[t + 1] - g[t];
        if (d < 0.0) d = 0.0;
        double exponent = -4.0 * ne * d * rho_rate;
        double rho = 1.0 - std::exp(exponent);
        if (rho < 0.0) rho = 0.0;
        if (rho > 1.0) rho = 1.0;
        double stay = (1.0 - rho) + rho / (2.0 * ne);
        Td[t] = static_cast<float>(std::log10(stay));
        double sw = rho / (2.0 * ne);
        To[t] = (sw < 1e-300) ? kLogZeroF : static_cast<float>(std::log10(sw));
    }
}

// Shared output step: A14 (posterior dosage) from a normalized probability
// vector p[k] and the state alleles.  AP0 = sum p_k [a0_k == 1],
// AP1 = sum p_k [a1_k == 1], DS = AP0 + AP1 (definition fixed from
// sampling.cu:80-100 and the ImputationResult posterior layout,
// imputer.cpp:427-445, 459-465).
static inline void dosage_from_probs(const double* p, const uint8_t* a0, const uint8_t* a1,
                                     uint32_t L, double* ds, double* ap0, double* ap1) {
    double s0 = 0.0, s1 = 0.0;
    for (uint32_t k = 0; k < L; ++k) {
        if (a0[k] == 1) s0 += p[k];
        if (a1[k] == 1) s1 += p[k];
    }
    *ap0 = s0; *ap1 = s1; *ds = s0 + s1;
}

// ---------------------------------------------------------------------------
// A6-A9 + A14, FP64, literal fold form.  Restates forward_pass_kernel
// (forward_backward.cu:41-142), backward_pass_kernel (:156-317) and the
// sampler's renormalisation (sampling.cu:80-100) in double precision:
//   forward: a'_m[j] = fold(lse2, -999, [alpha_{m-1}[i] + T[i][j]]_i) + E_m[j];
//            c_m = max_j a'; alpha_m = a' - c_m   (alpha_0 = E_0 - c_0)
//   backward: beta_{M-1} = 0;
//            beta_m[i] = fold(lse2, -999, [T[i][j] + E_{m+1}[j] + beta_{m+1}[j]]_j)
//   posterior: p'_m = alpha + beta; post = p' - fold(lse2, -999, p')
//   sampler: mx = fold(fmax, -999, post); q = exp(post - mx); p = q / sum q
// The reference replays alpha from sqrt(M) checkpoints with the same
// operations, which reproduces the forward values; this restatement keeps all
// alpha instead.  O(B*M*L^2).
//
// a0/a1: [M][L] raw state alleles; gl: [B][M][3]; Td/To: [M-1] per interval;
// per_marker = 0 reproduces the reference (interval 0 for every step).
// Outputs (any may be null): prob [B][M][L], logpost [B][M][L], ds [B][M],
// ap [B][M][2], loglik [B] (sum of forward scale factors plus the log mass of
// the last normalized alpha).
// ---------------------------------------------------------------------------
int or_fb_f64(const uint8_t* a0, const uint8_t* a1, const float* gl,
              const float* Td, const float* To, uint32_t B, uint32_t M, uint32_t L,
              int per_marker, double* prob, double* logpost, double* ds, double* ap,
              double* loglik) {
    if (M == 0 || L == 0) return -1;
    if (M >= 2 && (!Td || !To)) return -2;
    std::vector<double> alpha(size_t(M) * L), beta(size_t(M) * L), E(size_t(M) * L);
    std::vector<double> ap_(L), p(L);
    for (uint32_t s = 0; s < B; ++s) {
        const float* gls = gl + size_t(s) * M * 3;
        for (uint32_t m = 0; m < M; ++m)
            for (uint32_t k = 0; k < L; ++k)
                E[size_t(m) * L + k] = emission_lookup(gls + size_t(m) * 3,
                                                       a0[size_t(m) * L + k], a1[size_t(m) * L + k]);
        // forward
        double ll = 0.0;
        {
            double mx = E[0];
            for (uint32_t k = 1; k < L; ++k) mx = std::fmax(mx, E[k]);
            for (uint32_t k = 0; k < L; ++k) alpha[k] = E[k] - mx;
            ll += mx;
        }
        for (uint32_t m = 1; m < M; ++m) {
            uint32_t t = interval_for_step(m, per_marker);
            double td = Td[t], to = To[t];
            const double* ap_prev = &alpha[size_t(m - 1) * L];
            for (uint32_t j = 0; j < L; ++j) {
                double S = kLogZero;
                for (uint32_t i = 0; i < L; ++i) S = lse2(S, ap_prev[i] + (i == j ? td : to));
                ap_[j] = S + E[size_t(m) * L + j];
            }
            double mx = ap_[0];
            for (uint32_t j = 1; j < L; ++j) mx = std::fmax(mx, ap_[j]);
            for (uint32_t j = 0; j < L; ++j) alpha[size_t(m) * L + j] = ap_[j] - mx;
            ll += mx;
        }
        {
            double S = 0.0;
            for (uint32_t k = 0; k < L; ++k) S += std::exp(alpha[size_t(M - 1) * L + k]);
            ll += std::log(S);
        }
        if (loglik) loglik[s] = ll;
        // backward
        for (uint32_t k = 0; k < L; ++k) beta[size_t(M - 1) * L + k] = 0.0;
        for (int64_t m = int64_t(M) - 2; m >= 0; --m) {
            uint32_t t = interval_for_step(uint32_t(m + 1), per_marker);
            double td = Td[t], to = To[t];
            const double* bn = &beta[size_t(m + 1) * L];
            const double* En = &E[size_t(m + 1) * L];
            for (uint32_t i = 0; i < L; ++i) {
                double S = kLogZero;
                for (uint32_t j = 0; j < L; ++j) S = lse2(S, (i == j ? td : to) + En[j] + bn[j]);
                beta[size_t(m) * L + i] = S;
            }
        }
        // posterior + sampler normalisation + dosage
        for (uint32_t m = 0; m < M; ++m) {
            double Z = kLogZero;
            for (uint32_t k = 0; k < L; ++k) {
                p[k] = alpha[size_t(m) * L + k] + beta[size_t(m) * L + k];
                Z = lse2(Z, p[k]);
            }
            double mx = kLogZero;
            for (uint32_t k = 0; k < L; ++k) { p[k] -= Z; mx = std::fmax(mx, p[k]); }
            size_t o = (size_t(s) * M + m) * L;
            if (logpost) for (uint32_t k = 0; k < L; ++k) logpost[o + k] = p[k];
            double sum = 0.0;
            for (uint32_t k = 0; k < L; ++k) { p[k] = std::exp(p[k] - mx); sum += p[k]; }
            for (uint32_t k = 0; k < L; ++k) p[k] /= sum;
            if (prob) for (uint32_t k = 0; k < L; ++k) prob[o + k] = p[k];
            double d, x0, x1;
            dosage_from_probs(p.data(), a0 + size_t(m) * L, a1 + size_t(m) * L, L, &d, &x0, &x1);
            size_t so = size_t(s) * M + m;
            if (ds) ds[so] = d;
            if (ap) { ap[2 * so] = x0; ap[2 * so + 1] = x1; }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Windowed forward-backward, FP64 fold form: the reference's default path when
// window_size > 0 and M > window_size (imputer.cpp:965-971), restated window
// by window as run_windowed_forward_backward drives it (imputer.cpp:1044-1174)
// with forward_pass_windowed_kernel (forward_backward.cu:406-491) and
// backward_pass_windowed_kernel (:497-626):
//   forward, windows in order: window w starts from the previous window's
//     final alpha (w > 0; :432-438, so marker w*W gets neither a transition
//     nor its emission) or from E at marker 0, rescales by the max, steps to
//     its end; each window's final alpha is kept (imputer.cpp:1101-1106);
//   backward, windows in reverse: the window's forward is re-run from that
//     window's OWN final alpha (w > 0; imputer.cpp:1120-1142), beta starts at
//     the next window's final beta or 0 (:530-533, so the step from marker
//     (w+1)*W is not taken), walks down the window forming alpha + beta, and
//     the window's posteriors are normalized by their log-sum-exp (:602-625).
// The sampler normalisation and dosage follow as in or_fb_f64.  loglik (not
// produced by the reference) is the sum of the first forward run's rescale
// factors plus the log mass of its final alpha.  W = 0 or M <= W is the
// non-windowed or_fb_f64.
// ---------------------------------------------------------------------------
int or_fb_f64_windowed(const uint8_t* a0, const uint8_t* a1, const float* gl,
                       const float* Td, const float* To, uint32_t B, uint32_t M, uint32_t L,
                       int per_marker, uint32_t W, double* prob, double* ds, double* ap,
                       double* loglik) {
    if (W == 0 || M <= W)
        return or_fb_f64(a0, a1, gl, Td, To, B, M, L, per_marker, prob, nullptr, ds, ap, loglik);
    if (L == 0 || !Td || !To) return -1;
    const uint32_t nw = (M + W - 1) / W;
    std::vector<double> E(size_t(M) * L), win(size_t(W) * L), fa(size_t(nw) * L);
    std::vector<double> post(size_t(M) * L), bc(L), bp(L), x(L), p(L);
    // one window's forward (:406-491) into win[0..n-1]; returns the rescale sum
    auto fwd_window = [&](uint32_t start, uint32_t n, const double* init) {
        double ll = 0.0;
        for (uint32_t k = 0; k < L; ++k) win[k] = init ? init[k] : E[size_t(start) * L + k];
        double mx = win[0];
        for (uint32_t k = 1; k < L; ++k) mx = std::fmax(mx, win[k]);
        for (uint32_t k = 0; k < L; ++k) win[k] -= mx;
        ll += mx;
        for (uint32_t w = 1; w < n; ++w) {
            const uint32_t m = start + w, t = interval_for_step(m, per_marker);
            const double td = Td[t], to = To[t];
            const double* prev = &win[size_t(w - 1) * L];
            double* cur = &win[size_t(w) * L];
            for (uint32_t j = 0; j < L; ++j) {
                double S = kLogZero;
                for (uint32_t i = 0; i < L; ++i) S = lse2(S, prev[i] + (i == j ? td : to));
                cur[j] = S + E[size_t(m) * L + j];
            }
            mx = cur[0];
            for (uint32_t j = 1; j < L; ++j) mx = std::fmax(mx, cur[j]);
            for (uint32_t j = 0; j < L; ++j) cur[j] -= mx;
            ll += mx;
        }
        return ll;
    };
    for (uint32_t s = 0; s < B; ++s) {
        const float* gls = gl + size_t(s) * M * 3;
        for (uint32_t m = 0; m < M; ++m)
            for (uint32_t k = 0; k < L; ++k)
                E[size_t(m) * L + k] = emission_lookup(gls + size_t(m) * 3, a0[size_t(m) * L + k],
                                                       a1[size_t(m) * L + k]);
        double ll = 0.0;
        for (uint32_t w = 0; w < nw; ++w) {
            const uint32_t start = w * W, n = std::min(W, M - start);
            ll += fwd_window(start, n, w > 0 ? &fa[size_t(w - 1) * L] : nullptr);
            std::copy(&win[size_t(n - 1) * L], &win[size_t(n) * L], &fa[size_t(w) * L]);
        }
        {
            double S = 0.0;
            for (uint32_t k = 0; k < L; ++k) S += std::exp(fa[size_t(nw - 1) * L + k]);
            ll += std::log(S);
        }
        if (loglik) loglik[s] = ll;
        for (int w = int(nw) - 1; w >= 0; --w) {
            const uint32_t start = uint32_t(w) * W, n = std::min(W, M - start);
            fwd_window(start, n, w > 0 ? &fa[size_t(w) * L] : nullptr);
            if (w == int(nw) - 1) std::fill(bc.begin(), bc.end(), 0.0);  // else: the next window's
            for (uint32_t k = 0; k < L; ++k)
                post[size_t(start + n - 1) * L + k] = win[size_t(n - 1) * L + k] + bc[k];
            for (int ww = int(n) - 2; ww >= 0; --ww) {
                const uint32_t m = start + uint32_t(ww), t = interval_for_step(m + 1, per_marker);
                const double td = Td[t], to = To[t];
                bp = bc;
                const double* En = &E[size_t(m + 1) * L];
                for (uint32_t i = 0; i < L; ++i) {
                    double S = kLogZero;
                    for (uint32_t j = 0; j < L; ++j) S = lse2(S, (i == j ? td : to) + En[j] + bp[j]);
                    bc[i] = S;
                }
                for (uint32_t k = 0; k < L; ++k)
                    post[size_t(m) * L + k] = win[size_t(ww) * L + k] + bc[k];
            }
            // bc now holds beta at the window's first marker: the next (lower)
            // window starts from it
        }
        for (uint32_t m = 0; m < M; ++m) {
            double Z = kLogZero;
            for (uint32_t k = 0; k < L; ++k) Z = lse2(Z, post[size_t(m) * L + k]);
            double mx = kLogZero;
            for (uint32_t k = 0; k < L; ++k) { x[k] = post[size_t(m) * L + k] - Z; mx = std::fmax(mx, x[k]); }
            double sum = 0.0;
            for (uint32_t k = 0; k < L; ++k) { p[k] = std::exp(x[k] - mx); sum += p[k]; }
            for (uint32_t k = 0; k < L; ++k) p[k] /= sum;
            const size_t o = (size_t(s) * M + m) * L;
            if (prob) for (uint32_t k = 0; k < L; ++k) prob[o + k] = p[k];
            double d, x0, x1;
            dosage_from_probs(p.data(), a0 + size_t(m) * L, a1 + size_t(m) * L, L, &d, &x0, &x1);
            const size_t so = size_t(s) * M + m;
            if (ds) ds[so] = d;
            if (ap) { ap[2 * so] = x0; ap[2 * so + 1] = x1; }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Same quantities as or_fb_f64, FP64, using the rank-1-plus-diagonal structure
// of the Li-Stephens transition (O(L) per marker instead of O(L^2)):
//   sum_i e^{alpha_i + T_ij} = e^{To} sum_i e^{alpha_i} + (e^{Td} - e^{To}) e^{alpha_j}
// so fold(lse2, -999, ...) = LSE(-999, To + LSE(alpha), U + alpha_j) with
// U = log(e^{Td} - e^{To}).  Used where the L^2 fold is too slow; pinned
// against or_fb_f64 by the tests.  Threads over samples.
// ---------------------------------------------------------------------------
static inline double lse3(double a, double b, double c) {
    double mx = std::fmax(a, std::fmax(b, c));
    return mx + std::log(std::exp(a - mx) + std::exp(b - mx) + std::exp(c - mx));
}

static void fb_fast_one(const uint8_t* a0, const uint8_t* a1, const float* gls,
                        const double* TdU, const double* ToU, uint32_t M, uint32_t L,
                        int per_marker, double* prob, double* ds, double* ap, double* loglik,
                        std::vector<double>& alpha, std::vector<double>& E,
                        std::vector<double>& bcur, std::vector<double>& bnext,
                        std::vector<double>& p) {
    for (uint32_t m = 0; m < M; ++m)
        for (uint32_t k = 0; k < L; ++k)
            E[size_t(m) * L + k] = emission_lookup(gls + size_t(m) * 3, a0[size_t(m) * L + k],
                                                   a1[size_t(m) * L + k]);
    double ll = 0.0;
    {
        double mx = E[0];
        for (uint32_t k = 1; k < L; ++k) mx = std::fmax(mx, E[k]);
        for (uint32_t k = 0; k < L; ++k) alpha[k] = E[k] - mx;
        ll += mx;
    }
    for (uint32_t m = 1; m < M; ++m) {
        uint32_t t = interval_for_step(m, per_marker);
        const double* prev = &alpha[size_t(m - 1) * L];
        double mxp = prev[0];
        for (uint32_t i = 1; i < L; ++i) mxp = std::fmax(mxp, prev[i]);
        double sA = 0.0;
        for (uint32_t i = 0; i < L; ++i) sA += std::exp(prev[i] - mxp);
        double lA = mxp + std::log(sA);
        double* cur = &alpha[size_t(m) * L];
        double mx = -INFINITY;
        for (uint32_t j = 0; j < L; ++j) {
            cur[j] = lse3(kLogZero, ToU[2 * t] + lA, ToU[2 * t + 1] + prev[j]) + E[size_t(m) * L + j];
            mx = std::fmax(mx, cur[j]);
        }
        for (uint32_t j = 0; j < L; ++j) cur[j] -= mx;
        ll += mx;
    }
    {
        double S = 0.0;
        for (uint32_t k = 0; k < L; ++k) S += std::exp(alpha[size_t(M - 1) * L + k]);
        ll += std::log(S);
    }
    if (loglik) *loglik = ll;
    (void)TdU;
    // backward + posterior, marker by marker from the end
    for (uint32_t k = 0; k < L; ++k) bcur[k] = 0.0;
    for (int64_t m = int64_t(M) - 1; m >= 0; --m) {
        if (m < int64_t(M) - 1) {
            uint32_t t = interval_for_step(uint32_t(m + 1), per_marker);
            const double* En = &E[size_t(m + 1) * L];
            double mxw = -INFINITY;
            for (uint32_t j = 0; j < L; ++j) mxw = std::fmax(mxw, En[j] + bnext[j]);
            double sW = 0.0;
            for (uint32_t j = 0; j < L; ++j) sW += std::exp(En[j] + bnext[j] - mxw);
            double lW = mxw + std::log(sW);
            for (uint32_t i = 0; i < L; ++i)
                bcur[i] = lse3(kLogZero, ToU[2 * t] + lW, ToU[2 * t + 1] + En[i] + bnext[i]);
        }
        const double* al = &alpha[size_t(m) * L];
        double mx = -INFINITY;
        for (uint32_t k = 0; k < L; ++k) { p[k] = al[k] + bcur[k]; mx = std::fmax(mx, p[k]); }
        double sum = 0.0;
        for (uint32_t k = 0; k < L; ++k) { p[k] = std::exp(p[k] - mx); sum += p[k]; }
        for (uint32_t k = 0; k < L; ++k) p[k] /= sum;
        if (prob) for (uint32_t k = 0; k < L; ++k) prob[size_t(m) * L + k] = p[k];
        double d, x0, x1;
        dosage_from_probs(p.data(), a0 + size_t(m) * L, a1 + size_t(m) * L, L, &d, &x0, &x1);
        if (ds) ds[m] = d;
        if (ap) { ap[2 * size_t(m)] = x0; ap[2 * size_t(m) + 1] = x1; }
        std::swap(bcur, bnext);
    }
}

int or_fb_f64_fast(const uint8_t* a0, const uint8_t* a1, const float* gl,
                   const float* Td, const float* To, uint32_t B, uint32_t M, uint32_t L,
                   int per_marker, double* prob, double* ds, double* ap, double* loglik,
                   int nthreads) {
    if (M == 0 || L == 0) return -1;
    if (M >= 2 && (!Td || !To)) return -2;
    // Per interval: To (log) and U = log(e^{Td} - e^{To}) = Td + log1p(-e^{To-Td}).
    uint32_t nint = M >= 2 ? M - 1 : 0;
    std::vector<double> ToU(2 * size_t(std::max<uint32_t>(nint, 1)), 0.0);
    for (uint32_t t = 0; t < nint; ++t) {
        double td = Td[t], to = To[t];
        ToU[2 * t] = to;
        double r = std::exp(to - td);
        ToU[2 * t + 1] = (r >= 1.0) ? -INFINITY : td + std::log1p(-r);
    }
    if (nthreads < 1) nthreads = 1;
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        std::vector<double> alpha(size_t(M) * L), E(size_t(M) * L), bc(L), bn(L), p(L);
        for (;;) {
            uint32_t s = next.fetch_add(1);
            if (s >= B) break;
            fb_fast_one(a0, a1, gl + size_t(s) * M * 3, nullptr, ToU.data(), M, L, per_marker,
                        prob ? prob + size_t(s) * M * L : nullptr, ds ? ds + size_t(s) * M : nullptr,
                        ap ? ap + size_t(s) * M * 2 : nullptr, loglik ? loglik + s : nullptr,
                        alpha, E, bc, bn, p);
        }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < nthreads; ++i) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    return 0;
}

// ---------------------------------------------------------------------------
// FP32 "faithful" restatement: the reference's operation order in float with
// libm expf/logf (forward_backward.cu:41-317, sampling.cu:77-100).  Not an
// oracle for tolerances: it exists to quantify how far plain FP32 log-space
// arithmetic (with beta resting on the -999 floor) drifts from FP64.
// Output: prob [B][M][L] float (the sampler's normalized vector), ds [B][M].
// ---------------------------------------------------------------------------
int or_fb_f32_faithful(const uint8_t* a0, const uint8_t* a1, const float* gl,
                       const float* Td, const float* To, uint32_t B, uint32_t M, uint32_t L,
                       int per_marker, float* prob, float* ds) {
    if (M == 0 || L == 0) return -1;
    std::vector<float> alpha(size_t(M) * L), beta(size_t(M) * L), E(size_t(M) * L), ap_(L), p(L);
    for (uint32_t s = 0; s < B; ++s) {
        const float* gls = gl + size_t(s) * M * 3;
        for (uint32_t m = 0; m < M; ++m)
            for (uint32_t k = 0; k < L; ++k)
                E[size_t(m) * L + k] = emission_lookup(gls + size_t(m) * 3, a0[size_t(m) * L + k],
                                                       a1[size_t(m) * L + k]);
        for (uint32_t k = 0; k < L; ++k) alpha[k] = E[k];
        {
            float mx = alpha[0];
            for (uint32_t k = 1; k < L; ++k) mx = std::fmax(mx, alpha[k]);
            for (uint32_t k = 0; k < L; ++k) alpha[k] -= mx;
        }
        for (uint32_t m = 1; m < M; ++m) {
            uint32_t t = interval_for_step(m, per_marker);
            for (uint32_t j = 0; j < L; ++j) {
                float S = kLogZeroF;
                for (uint32_t i = 0; i < L; ++i) {
                    float tr = (i == j) ? Td[t] : To[t];
                    float val = alpha[size_t(m - 1) * L + i] + tr;
                    S = lse2f(S, val);
                }
                ap_[j] = S + E[size_t(m) * L + j];
            }
            float mx = ap_[0];
            for (uint32_t j = 1; j < L; ++j) mx = std::fmax(mx, ap_[j]);
            for (uint32_t j = 0; j < L; ++j) alpha[size_t(m) * L + j] = ap_[j] - mx;
        }
        for (uint32_t k = 0; k < L; ++k) beta[size_t(M - 1) * L + k] = 0.0f;
        for (int64_t m = int64_t(M) - 2; m >= 0; --m) {
            uint32_t t = interval_for_step(uint32_t(m + 1), per_marker);
            for (uint32_t i = 0; i < L; ++i) {
                float S = kLogZeroF;
                for (uint32_t j = 0; j < L; ++j) {
                    float tr = (i == j) ? Td[t] : To[t];
                    float val = tr + E[size_t(m + 1) * L + j] + beta[size_t(m + 1) * L + j];
                    S = lse2f(S, val);
                }
                beta[size_t(m) * L + i] = S;
            }
        }
        for (uint32_t m = 0; m < M; ++m) {
            float Z = kLogZeroF;
            for (uint32_t k = 0; k < L; ++k) {
                p[k] = alpha[size_t(m) * L + k] + beta[size_t(m) * L + k];
                Z = lse2f(Z, p[k]);
            }
            float mx = kLogZeroF;
            for (uint32_t k = 0; k < L; ++k) { p[k] = p[k] - Z; mx = std::fmax(mx, p[k]); }
            float cs = 0.0f;
            for (uint32_t k = 0; k < L; ++k) { p[k] = std::exp(p[k] - mx); cs += p[k]; }
            for (uint32_t k = 0; k < L; ++k) p[k] /= cs;
            size_t o = (size_t(s) * M + m) * L;
            if (prob) for (uint32_t k = 0; k < L; ++k) prob[o + k] = p[k];
            if (ds) {
                float d = 0.0f;
                for (uint32_t k = 0; k < L; ++k)
                    d += p[k] * float((a0[size_t(m) * L + k] == 1) + (a1[size_t(m) * L + k] == 1));
                ds[size_t(s) * M + m] = d;
            }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): ten rounds, key bumped by the Weyl constants
// between rounds.  ctr[4], key[2] -> out[4].
// ---------------------------------------------------------------------------
void or_philox4x32_10(const uint32_t* ctr_in, const uint32_t* key_in, uint32_t* out) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// The sampler's uniform draw for (job sample s, marker m): output word (m & 3)
// of Philox4x32-10 with counter (m >> 2, s, stream, 0) and key (seed lo,
// seed hi), mapped to [0, 1) as (x >> 8) * 2^-24 (exact in FP32).  One
// Philox call serves four consecutive markers.
float or_sample_uniform(uint64_t seed, uint32_t s, uint32_t m, uint32_t stream) {
    uint32_t ctr[4] = {m >> 2, s, stream, 0u};
    uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    uint32_t out[4];
    or_philox4x32_10(ctr, key, out);
    return float(out[m & 3u] >> 8) * (1.0f / 16777216.0f);
}

// Inverse-CDF draw on an FP32 probability vector, the reference's loop
// (sampling.cu:104-113): sequential FP32 cumulative sum, first k with
// u < cumsum_k, else state 0.
uint32_t or_sample_from_probs(const float* p, uint32_t L, float u) {
    float cs = 0.0f;
    for (uint32_t k = 0; k < L; ++k) {
        cs += p[k];
        if (u < cs) return k;
    }
    return 0;
}

// Argmax, first maximum (sampling.cu:159-169).
uint32_t or_argmax(const float* p, uint32_t L) {
    uint32_t best = 0;
    float bp = p[0];
    for (uint32_t k = 1; k < L; ++k) if (p[k] > bp) { bp = p[k]; best = k; }
    return best;
}

// Host replay of the sampler over a whole job, given the device's FP32
// probability vectors prob [B][M][L] and raw state alleles a0/a1 [M][L]:
// hap [B][2][M] (the ImputationResult layout, imputer.hpp:191).
// mode 0 = Philox draw, 1 = argmax; sample_base is the global index of
// sample 0 of the job.
void or_sample_replay(const float* prob, const uint8_t* a0, const uint8_t* a1,
                      uint32_t B, uint32_t M, uint32_t L, uint64_t seed, uint32_t stream,
                      int mode, uint32_t sample_base, uint8_t* hap) {
    for (uint32_t s = 0; s < B; ++s) {
        for (uint32_t m = 0; m < M; ++m) {
            const float* p = prob + (size_t(s) * M + m) * L;
            uint32_t k = mode == 1 ? or_argmax(p, L)
                                   : or_sample_from_probs(p, L, or_sample_uniform(seed, sample_base + s, m, stream));
            hap[size_t(s) * 2 * M + m] = a0[size_t(m) * L + k];
            hap[size_t(s) * 2 * M + M + m] = a1[size_t(m) * L + k];
        }
    }
}

// ---------------------------------------------------------------------------
// The reference CPU path, timed as the CPU baseline: PBWT sweep
// (pbwt_index.cpp:153-208), per-(sample, marker) std::sort selection
// (imputer.cpp:902-928 -> pbwt_index.cpp:20-44), pairing, emission,
// FP32 forward with floor(sqrt(M)) checkpoints and the backward pass that
// replays alpha from the last checkpoint for every marker
// (forward_backward.cu:41-317), then the inverse-CDF draw (sampling.cu:77-127)
// with the Philox uniform in place of cuRAND.  Samples are spread over
// nthreads std::threads.  hap_out [B][2][M].  Returns 0.
// ---------------------------------------------------------------------------
int or_cpu_reference_impute(const uint8_t* ref, const float* gl, const float* Td,
                            const float* To, uint32_t M, uint32_t H, uint32_t B, uint32_t L,
                            uint64_t seed, int nthreads, uint8_t* hap_out) {
    if (M == 0 || H == 0 || L == 0) return -1;
    // index build (single thread, as the reference)
    std::vector<uint32_t> prefix(size_t(M) * H), div(size_t(M) * H);
    or_pbwt_build(ref, M, H, prefix.data(), div.data());
    uint32_t C = uint32_t(std::sqrt(float(M)));
    if (C < 1) C = 1;
    uint32_t ncp = (M + C - 1) / C;
    if (nthreads < 1) nthreads = 1;
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        std::vector<uint32_t> st(size_t(M) * L), best(L);
        std::vector<uint8_t> a0(size_t(M) * L), a1(size_t(M) * L);
        std::vector<float> E(size_t(M) * L), cp(size_t(ncp) * L), sc(M), post(size_t(M) * L);
        std::vector<float> cur(L), prv(L), ar(L), tmp(L), bc(L), bp(L);
        for (;;) {
            uint32_t s = next.fetch_add(1);
            if (s >= B) break;
            for (uint32_t m = 0; m < M; ++m) {
                or_select_states_sort(&prefix[size_t(m) * H], &div[size_t(m) * H], H, L, best.data());
                for (uint32_t k = 0; k < L; ++k) {
                    uint32_t h0 = best[k], h1 = (h0 + 1) % H;
                    a0[size_t(m) * L + k] = ref[size_t(m) * H + h0];
                    a1[size_t(m) * L + k] = ref[size_t(m) * H + h1];
                }
            }
            const float* gls = gl + size_t(s) * M * 3;
            for (uint32_t m = 0; m < M; ++m)
                for (uint32_t k = 0; k < L; ++k)
                    E[size_t(m) * L + k] = emission_lookup(gls + size_t(m) * 3, a0[size_t(m) * L + k],
                                                           a1[size_t(m) * L + k]);
            const float td = M > 1 ? Td[0] : 0.f, to = M > 1 ? To[0] : 0.f;
            auto step = [&](const float* in, float* out, uint32_t m) {
                for (uint32_t j = 0; j < L; ++j) {
                    float S = kLogZeroF;
                    for (uint32_t i = 0; i < L; ++i) S = lse2f(S, in[i] + (i == j ? td : to));
                    out[j] = S + E[size_t(m) * L + j];
                }
            };
            // forward with checkpoints
            for (uint32_t k = 0; k < L; ++k) cur[k] = E[k];
            {
                float mx = cur[0];
                for (uint32_t k = 1; k < L; ++k) mx = std::fmax(mx, cur[k]);
                sc[0] = mx;
                for (uint32_t k = 0; k < L; ++k) cur[k] -= mx;
                for (uint32_t k = 0; k < L; ++k) cp[k] = cur[k];
            }
            for (uint32_t m = 1; m < M; ++m) {
                std::swap(cur, prv);
                step(prv.data(), cur.data(), m);
                float mx = cur[0];
                for (uint32_t k = 1; k < L; ++k) mx = std::fmax(mx, cur[k]);
                sc[m] = mx;
                for (uint32_t k = 0; k < L; ++k) cur[k] -= mx;
                if (m % C == 0) for (uint32_t k = 0; k < L; ++k) cp[size_t(m / C) * L + k] = cur[k];
            }
            // backward with per-marker replay from the last checkpoint
            auto replay = [&](uint32_t m) {
                uint32_t c = m / C;
                for (uint32_t k = 0; k < L; ++k) ar[k] = cp[size_t(c) * L + k];
                for (uint32_t mm = c * C + 1; mm <= m; ++mm) {
                    step(ar.data(), tmp.data(), mm);
                    for (uint32_t k = 0; k < L; ++k) ar[k] = tmp[k] - sc[mm];
                }
            };
            for (uint32_t k = 0; k < L; ++k) bc[k] = 0.0f;
            replay(M - 1);
            for (uint32_t k = 0; k < L; ++k) post[size_t(M - 1) * L + k] = ar[k] + bc[k];
            for (int64_t m = int64_t(M) - 2; m >= 0; --m) {
                std::swap(bc, bp);
                for (uint32_t i = 0; i < L; ++i) {
                    float S = kLogZeroF;
                    for (uint32_t j = 0; j < L; ++j)
                        S = lse2f(S, (i == j ? td : to) + E[size_t(m + 1) * L + j] + bp[j]);
                    bc[i] = S;
                }
                replay(uint32_t(m));
                for (uint32_t k = 0; k < L; ++k) post[size_t(m) * L + k] = ar[k] + bc[k];
            }
            for (uint32_t m = 0; m < M; ++m) {
                float Z = kLogZeroF;
                for (uint32_t k = 0; k < L; ++k) Z = lse2f(Z, post[size_t(m) * L + k]);
                float mx = kLogZeroF;
                for (uint32_t k = 0; k < L; ++k) {
                    tmp[k] = post[size_t(m) * L + k] - Z;
                    mx = std::fmax(mx, tmp[k]);
                }
                float cs = 0.0f;
                for (uint32_t k = 0; k < L; ++k) { tmp[k] = std::exp(tmp[k] - mx); cs += tmp[k]; }
                for (uint32_t k = 0; k < L; ++k) tmp[k] /= cs;
                uint32_t k = or_sample_from_probs(tmp.data(), L, or_sample_uniform(seed, s, m, 0));
                hap_out[size_t(s) * 2 * M + m] = a0[size_t(m) * L + k];
                hap_out[size_t(s) * 2 * M + M + m] = a1[size_t(m) * L + k];
            }
        }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < nthreads; ++i) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    return 0;
}

}  // extern "C"
