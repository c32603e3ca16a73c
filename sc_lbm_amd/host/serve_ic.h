// serve_ic.h -- seeded initial densities of generated served jobs.  Shared
// by the host (generator, CPU checks) and the device initialiser, which must
// evaluate the identical double-precision expression site by site.
//
// Per-site noise: u = the top 53 bits of the SplitMix64 output for counter
// k = y*nx + x from the component's seed, as a double in [0, 1); density
// rho * (1 + (2u - 1) * noise).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SCL_HD __host__ __device__
#else
#define SCL_HD
#endif

namespace scl {
namespace serve {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

// SplitMix64 output number k (k = 0, 1, ...) of the stream seeded by `seed`.
SCL_HD inline uint64_t splitmix64_at(uint64_t seed, uint64_t k) {
    uint64_t z = seed + (k + 1) * kGolden;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

SCL_HD inline double uniform01_at(uint64_t seed, uint64_t k) {
    return (double)(splitmix64_at(seed, k) >> 11) * (1.0 / 9007199254740992.0);
}

struct IcParams {
    int nx, ny, ic;
    double rho_hi, rho_lo, radius, noise;
    uint64_t seed;   // component A; B uses seed + 1
};

// ic: 0 uniform, 1 droplet (disk about (nx/2, ny/2)), 2 stratified (rows y < ny/2).
SCL_HD inline double ic_density_at(const IcParams& p, int c, int x, int y) {
    bool a_hi = true;
    if (p.ic == 1) {
        const long long dx = x - p.nx / 2, dy = y - p.ny / 2;
        a_hi = (double)(dx * dx + dy * dy) <= p.radius * p.radius;
    } else if (p.ic == 2) {
        a_hi = y < p.ny / 2;
    }
    double rho = ((c == 0) == a_hi) ? p.rho_hi : p.rho_lo;
    if (p.noise != 0.0) {
        const double u = uniform01_at(p.seed + (uint64_t)c, (uint64_t)y * (uint64_t)p.nx + (uint64_t)x);
        rho = rho * (1.0 + (2.0 * u - 1.0) * p.noise);
    }
    return rho;
}

}  // namespace serve
}  // namespace scl
