// Seeded synthetic panels and targets (declared in include/swiftimpute/synth.h).
//
// Panel: F founders with per-marker allele frequency ~ Beta(1/2, 1/2)
// (sampled exactly as sin^2(pi U / 2)); each panel haplotype copies founders,
// switching at the events of a rate-1-per-Morgan Poisson process along the
// genetic map to a uniformly drawn founder; each allele flips with the
// mutation probability.  Targets: two haplotypes per sample copying panel
// haplotypes by the same process, then adjusted toward a per-sample
// heterozygosity target, then typed at a regular subset of markers with
// genotyping error.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../../include/swiftimpute/synth.h"
#include "philox_host.hpp"

using si_host::Stream;

namespace {

enum Tag : uint32_t {
    kFreq = 1,
    kFounder = 2,
    kSpacing = 3,
    kHot = 4,
    kSwitch = 5,
    kMut = 6,
    kTgtSwitch = 7,
    kTgtHet = 8,
    kTgtErr = 9
};

template <class F>
void parallel_for(uint32_t n, int nthreads, F&& f) {
    if (nthreads < 1) nthreads = 1;
    if (uint32_t(nthreads) > n) nthreads = int(std::max<uint32_t>(n, 1));
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= n) break;
            f(i);
        }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < nthreads; ++i) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

// Copying path along the map: marker index where the next switch takes
// effect (first marker whose genetic position exceeds the event).
struct Path {
    Stream rng;
    const double* G;
    uint32_t M;
    double x = 0.0;
    Path(uint64_t seed, uint32_t tag, uint32_t id, const double* G_, uint32_t M_)
        : rng(seed, tag, id, 0), G(G_), M(M_) {
        x = G[0];
    }
    uint32_t next_switch() {
        x += -std::log(rng.uniform());
        const double* it = std::upper_bound(G, G + M, x);
        return uint32_t(it - G);  // M if beyond the last marker
    }
};

}  // namespace

extern "C" {

void si_synth_panel_defaults(si_synth_panel_cfg* c, uint32_t H, uint32_t M, uint64_t seed) {
    c->n_haplotypes = H;
    c->n_markers = M;
    c->n_founders = 40;
    c->spacing_mode = 0;
    c->seed = seed;
    c->mutation = 1e-3;
    c->spacing_bp = 50000.0;
    c->lognorm_median_bp = 20000.0;
    c->lognorm_sigma = 1.5;
    c->hotspot_frac = 0.01;
    c->hotspot_mult = 50.0;
}

int si_synth_panel(const si_synth_panel_cfg* cfg, uint64_t* cols, uint64_t* pos_out, double* cM_out,
                   double* morgans_out, int n_threads) {
    if (!cfg || !cols) return -1;
    const uint32_t H = cfg->n_haplotypes, M = cfg->n_markers, F = cfg->n_founders;
    if (H == 0 || M == 0 || F == 0 || F > 64) return -1;
    const uint64_t seed = cfg->seed;
    const uint32_t W = (H + 63) / 64;

    // positions and genetic map
    std::vector<uint64_t> pos(M);
    std::vector<double> G(M);  // Morgans
    if (cfg->spacing_mode == 0) {
        for (uint32_t m = 0; m < M; ++m) pos[m] = uint64_t(cfg->spacing_bp * (m + 1));
        for (uint32_t m = 0; m < M; ++m) G[m] = double(pos[m]) / 1e8;
    } else {
        Stream sp(seed, kSpacing, 0, 0), hot(seed, kHot, 0, 0);
        const double mu = std::log(cfg->lognorm_median_bp);
        uint64_t p = uint64_t(cfg->lognorm_median_bp);
        G[0] = double(p) / 1e8;
        pos[0] = p;
        for (uint32_t m = 1; m < M; ++m) {
            double gap = std::exp(mu + cfg->lognorm_sigma * sp.normal());
            uint64_t ig = uint64_t(std::llround(gap));
            if (ig < 1) ig = 1;
            p += ig;
            pos[m] = p;
            const double mult = hot.uniform() < cfg->hotspot_frac ? cfg->hotspot_mult : 1.0;
            G[m] = G[m - 1] + double(ig) / 1e8 * mult;
        }
    }
    if (pos_out) std::memcpy(pos_out, pos.data(), sizeof(uint64_t) * M);
    if (cM_out)
        for (uint32_t m = 0; m < M; ++m) cM_out[m] = G[m] * 100.0;
    if (morgans_out) std::memcpy(morgans_out, G.data(), sizeof(double) * M);

    // founder alleles, bit f of FB[m]
    std::vector<uint64_t> FB(M, 0);
    parallel_for(M, n_threads, [&](uint32_t m) {
        Stream fr(seed, kFreq, m, 0), fa(seed, kFounder, m, 0);
        const double s = std::sin(1.5707963267948966 * fr.uniform());
        const double f = s * s;
        uint64_t w = 0;
        for (uint32_t k = 0; k < F; ++k)
            if (fa.uniform() < f) w |= 1ull << k;
        FB[m] = w;
    });

    const double lq = std::log1p(-cfg->mutation);
    parallel_for(W, n_threads, [&](uint32_t w) {
        const uint32_t h0 = w * 64, nb = std::min<uint32_t>(64, H - h0);
        std::vector<Path> paths;
        paths.reserve(nb);
        uint32_t cur[64], nsw[64];
        uint64_t nmut[64];
        std::vector<Stream> fnd, mut;
        fnd.reserve(nb);
        mut.reserve(nb);
        for (uint32_t b = 0; b < nb; ++b) {
            paths.emplace_back(seed, kSwitch, h0 + b, G.data(), M);
            fnd.emplace_back(seed, kSwitch, h0 + b, 1);
            mut.emplace_back(seed, kMut, h0 + b, 0);
            cur[b] = fnd[b].below(F);
            nsw[b] = paths[b].next_switch();
            nmut[b] = (cfg->mutation > 0.0)
                          ? uint64_t(std::floor(std::log(mut[b].uniform()) / lq))
                          : ~0ull;
        }
        for (uint32_t m = 0; m < M; ++m) {
            uint64_t word = 0;
            for (uint32_t b = 0; b < nb; ++b) {
                while (nsw[b] <= m) {  // a switch effective at or before m
                    cur[b] = fnd[b].below(F);
                    nsw[b] = std::max(paths[b].next_switch(), nsw[b]);
                    if (nsw[b] <= m) continue;
                }
                uint64_t bit = (FB[m] >> cur[b]) & 1ull;
                if (nmut[b] == m) {
                    bit ^= 1ull;
                    nmut[b] += 1 + uint64_t(std::floor(std::log(mut[b].uniform()) / lq));
                }
                word |= bit << b;
            }
            cols[size_t(m) * W + w] = word;
        }
    });
    return 0;
}

void si_synth_target_defaults(si_synth_target_cfg* c, uint32_t B, uint32_t M, uint32_t typed_every,
                              uint64_t seed) {
    c->n_samples = B;
    c->typed_every = typed_every;
    c->typed_offset = 0;
    c->code_pitch = (M + 15) / 16 * 16;
    c->seed = seed;
    c->het_lo = 0.12;
    c->het_hi = 0.40;
    c->geno_error = 0.005;
}

int si_synth_targets(const si_synth_target_cfg* cfg, const uint64_t* cols, uint32_t H, uint32_t M,
                     const double* G, uint8_t* codes, uint8_t* truth, int n_threads) {
    if (!cfg || !cols || !G || !codes || H == 0 || M == 0) return -1;
    if (cfg->code_pitch < M || cfg->typed_every == 0) return -1;
    const uint32_t W = (H + 63) / 64;
    const uint64_t seed = cfg->seed;
    auto allele = [&](uint32_t m, uint32_t h) -> uint8_t {
        return uint8_t((cols[size_t(m) * W + (h >> 6)] >> (h & 63)) & 1ull);
    };
    parallel_for(cfg->n_samples, n_threads, [&](uint32_t s) {
        std::vector<uint8_t> hp[2] = {std::vector<uint8_t>(M), std::vector<uint8_t>(M)};
        for (uint32_t k = 0; k < 2; ++k) {
            Path path(seed, kTgtSwitch, 2 * s + k, G, M);
            Stream src(seed, kTgtSwitch, 2 * s + k, 1);
            uint32_t cur = src.below(H), nsw = path.next_switch();
            for (uint32_t m = 0; m < M; ++m) {
                while (nsw <= m) {
                    cur = src.below(H);
                    nsw = std::max(path.next_switch(), nsw);
                    if (nsw <= m) continue;
                }
                hp[k][m] = allele(m, cur);
            }
        }
        // heterozygosity adjustment toward the sample's target
        Stream het(seed, kTgtHet, s, 0);
        const double target = cfg->het_lo + (cfg->het_hi - cfg->het_lo) * het.uniform();
        uint32_t nhet = 0;
        for (uint32_t m = 0; m < M; ++m) nhet += hp[0][m] != hp[1][m];
        const double nat = double(nhet) / M;
        if (target < nat) {
            const double lam = 1.0 - target / nat;
            for (uint32_t m = 0; m < M; ++m)
                if (hp[0][m] != hp[1][m] && het.uniform() < lam) hp[1][m] = hp[0][m];
        } else if (nat < 1.0) {
            const double lam = (target - nat) / (1.0 - nat);
            for (uint32_t m = 0; m < M; ++m)
                if (hp[0][m] == hp[1][m] && het.uniform() < lam) hp[1][m] = uint8_t(1 - hp[0][m]);
        }
        Stream err(seed, kTgtErr, s, 0);
        uint8_t* row = codes + size_t(s) * cfg->code_pitch;
        for (uint32_t m = 0; m < M; ++m) {
            uint8_t c = 3;
            if (m % cfg->typed_every == cfg->typed_offset % cfg->typed_every) {
                c = uint8_t(hp[0][m] + hp[1][m]);
                if (err.uniform() < cfg->geno_error) c = uint8_t((c + 1 + err.below(2)) % 3);
            }
            row[m] = c;
        }
        for (uint32_t m = M; m < cfg->code_pitch; ++m) row[m] = 3;
        if (truth) {
            std::memcpy(truth + size_t(s) * 2 * M, hp[0].data(), M);
            std::memcpy(truth + size_t(s) * 2 * M + M, hp[1].data(), M);
        }
    });
    return 0;
}

void si_hardcall_palette(float pal[12]) {
    const float z = -999.0f;
    const float v[12] = {0.f, z, z, z, 0.f, z, z, z, 0.f, 0.f, 0.f, 0.f};
    std::memcpy(pal, v, sizeof(v));
}

}  // extern "C"
