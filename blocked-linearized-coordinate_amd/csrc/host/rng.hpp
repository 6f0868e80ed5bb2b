// rng.hpp -- counter-based random numbers and the samplers of the synthetic
// tensor generator.  Every draw is a pure function of (seed, stream, counter),
// so results do not depend on thread count or scheduling.
#pragma once

#include <cmath>
#include <cstdint>

namespace blco {

// SplitMix64 finaliser (Steele, Lea, Flood 2014).
inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    uint64_t k = mix64(seed ^ mix64(stream * 0xD6E8FEB86659FD93ull));
    return mix64(k + ctr * 0x9E3779B97F4A7C15ull);
}

// Uniform double in [0, 1) with 53 random bits.
inline double rng_u01(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    return (double)(rng_u64(seed, stream, ctr) >> 11) * 0x1.0p-53;
}

// Random-stream ids.
enum : uint64_t {
    STREAM_INDEX = 0x100,    // + mode: coordinate draws
    STREAM_GEN_FACTOR = 0x200,  // + mode: Gamma(1,1) generating factors
    STREAM_VALUE = 0x300,    // Poisson values
    STREAM_INIT = 0x400      // + mode: CP initial factors
};

// Zero-truncated Poisson(lam) count (>= 1) drawn from the counter sequence
// (seed, STREAM_VALUE, base + j), j = 0, 1, ...  lam <= 30: inversion of the
// truncated CDF; lam > 30: PTRS transformed rejection (Hoermann 1993,
// "The transformed rejection method for generating Poisson random
// variables"), redrawn on 0.  Every multiply-add that is fused is written
// as std::fma and the file is compiled without contraction, so the GPU
// generator (the same steps in gen_dev.hip) forms the same doubles.
inline uint64_t ztpoisson(double lam, uint64_t seed, uint64_t base)
{
    uint64_t j = 0;
    if (!(lam > 0.0)) return 1;
    if (lam <= 30.0) {
        double p0 = std::exp(-lam);
        double u = rng_u01(seed, STREAM_VALUE, base + j++);
        double target = std::fma(u, 1.0 - p0, p0);  // P(K <= k) crossing, K >= 1
        double p = p0, c = p0;
        uint64_t k = 0;
        while (c < target && k < 10000) {
            ++k;
            p *= lam / (double)k;
            c += p;
        }
        return k < 1 ? 1 : k;
    }
    double slam = std::sqrt(lam), loglam = std::log(lam);
    double b = std::fma(slam, 2.53, 0.931);
    double a = std::fma(b, 0.02483, -0.059);
    double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        double U = rng_u01(seed, STREAM_VALUE, base + j++) - 0.5;
        double V = rng_u01(seed, STREAM_VALUE, base + j++);
        double us = 0.5 - std::fabs(U);
        double kf = std::floor(std::fma(2.0 * a / us + b, U, lam) + 0.43);
        if (us >= 0.07 && V <= vr) {
            if (kf >= 1.0) return (uint64_t)kf;
            continue;
        }
        if (kf < 0.0 || (us < 0.013 && V > us)) continue;
        if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
            std::fma(kf, loglam, -lam) - std::lgamma(kf + 1.0)) {
            if (kf >= 1.0) return (uint64_t)kf;
        }
    }
}

}  // namespace blco
