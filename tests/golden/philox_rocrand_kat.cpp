// Prints Philox4x32-10 outputs from rocRAND's open-source engine
// (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h) as JSON, for
// cross-checking the independent implementations in this repository.
// An engine seeded with (seed, subsequence, offset = 4 * k) evaluates the
// counter (k, 0, subsequence lo, subsequence hi) under key (seed lo, seed hi).
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdint>
#include <cstdio>

int main() {
    const unsigned long long seeds[] = {0ull, 12345ull, 0xa4093822299f31d0ull, 0xffffffffffffffffull};
    const unsigned long long subs[] = {0ull, 1ull, 7ull, 0x0000000500000003ull};
    const unsigned long long ks[] = {0ull, 1ull, 1000ull};
    std::printf("[\n");
    bool first = true;
    for (auto seed : seeds)
        for (auto sub : subs)
            for (auto k : ks) {
                rocrand_device::philox4x32_10_engine eng(seed, sub, 4 * k);
                unsigned w[4];
                for (int i = 0; i < 4; ++i) w[i] = eng.next();
                std::printf("%s{\"key\": [%u, %u], \"ctr\": [%u, %u, %u, %u], \"out\": [%u, %u, %u, %u]}",
                            first ? "" : ",\n", unsigned(seed), unsigned(seed >> 32), unsigned(k),
                            unsigned(k >> 32), unsigned(sub), unsigned(sub >> 32), w[0], w[1], w[2],
                            w[3]);
                first = false;
            }
    std::printf("\n]\n");
    return 0;
}
