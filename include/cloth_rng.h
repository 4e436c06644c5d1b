#pragma once
// SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): the one named generator behind every seeded
// workload of this library, so a seed fixes a workload on any platform.

#include <cmath>
#include <cstdint>

namespace cloth {

struct SplitMix64 {
    uint64_t state;
    explicit SplitMix64(uint64_t seed) : state(seed) {}

    uint64_t next_u64() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // Uniform double in [0, 1) from the top 53 bits.
    double uniform() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
    // Uniform double in [lo, hi).
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    // Log-uniform double in [lo, hi).
    double log_uniform(double lo, double hi) { return std::exp(uniform(std::log(lo), std::log(hi))); }
    // Uniform integer in [0, n).
    uint64_t below(uint64_t n) { return next_u64() % n; }
};

}  // namespace cloth
