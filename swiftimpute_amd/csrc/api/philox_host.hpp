// Host Philox4x32-10 (Salmon et al., SC'11) for counter-based, thread-order
// independent random streams in the synthetic data generator.
#pragma once

#include <cmath>
#include <cstdint>

namespace si_host {

struct U4 {
    uint32_t x, y, z, w;
};

inline U4 philox(U4 c, uint64_t seed) {
    uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = uint64_t(0xD2511F53u) * c.x, p1 = uint64_t(0xCD9E8D57u) * c.z;
        c = U4{uint32_t(p1 >> 32) ^ c.y ^ k0, uint32_t(p1), uint32_t(p0 >> 32) ^ c.w ^ k1,
               uint32_t(p0)};
    }
    return c;
}

// A sequential stream over one counter prefix (tag, a, b): draw i uses
// counter (tag, a, b, i / 4), word i % 4.
class Stream {
  public:
    Stream(uint64_t seed, uint32_t tag, uint32_t a, uint32_t b) : seed_(seed), tag_(tag), a_(a), b_(b) {}
    uint32_t next_u32() {
        if ((i_ & 3u) == 0) cur_ = philox(U4{tag_, a_, b_, i_ >> 2}, seed_);
        const uint32_t w[4] = {cur_.x, cur_.y, cur_.z, cur_.w};
        return w[(i_++) & 3u];
    }
    // uniform in (0, 1): 53-bit from two words, never 0
    double uniform() {
        const uint64_t hi = next_u32() >> 5, lo = next_u32() >> 6;
        return (double(hi * 67108864ull + lo) + 0.5) / 9007199254740992.0;
    }
    double normal() {  // Box-Muller
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    uint32_t below(uint32_t n) { return uint32_t(uniform() * n) % n; }

  private:
    uint64_t seed_;
    uint32_t tag_, a_, b_, i_ = 0;
    U4 cur_{0, 0, 0, 0};
};

}  // namespace si_host
