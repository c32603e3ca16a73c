/*
 * rng.h -- drop-in for the reference's RNG (include/rng.h:14-31 of
 * LBMFluids/sc_lbm): uniform doubles from std::mt19937.  The default
 * constructor seeds from std::random_device as the reference does; the
 * seeded constructor makes runs reproducible.
 */
#ifndef SCLBM_RNG_H
#define SCLBM_RNG_H

#include <cstdint>
#include <random>

class RNG {
public:
    RNG() : gen(std::random_device()()) {}
    explicit RNG(std::uint32_t seed) : gen(seed) {}

    /// Uniform in [dmin, dmax)
    double get_random(const double dmin, const double dmax) {
        std::uniform_real_distribution<double> dist(dmin, dmax);
        return dist(gen);
    }

private:
    std::mt19937 gen;
};

#endif
