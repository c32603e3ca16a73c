// dropin_bench.cpp -- the reference's phase-separation loop
// (benchmarks/two_phase_two_components_flows/phase_separation/phase_separation.cpp:80-116)
// written against the drop-in classes, timed.
//
// usage: dropin_bench <n> <steps> <precision 0|1|2> <read_every> [out.bin]
//   n x n periodic lattice, two components with randomly perturbed densities
//   (RNG seeded 2024; mass fraction 0.665 of rho_bar 2.0, G 0.9, tau 1), the
//   same in the deviation forms with rho0 = the component's mean density.
//   read_every k > 0: a host read of one population of each fluid after
//   every k-th step (each read issues the steps queued before it and
//   downloads the populations); k < 0: sclbm_flush() after every |k|-th
//   step (issues the queued steps, no download; -1 is one fused launch
//   sequence per step, as without deferral); 0: neither until the end.
//   One untimed step first, then `steps` timed ones.  Prints one JSON line;
//   with out.bin, writes both fluids' populations (raw doubles, A then B).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "geometry.h"
#include "lbm.h"
#include "rng.h"
#include "sc_lbm_amd.h"

int main(int argc, char** argv) {
    if (argc < 5) {
        std::cerr << "usage: dropin_bench <n> <steps> <precision 0|1|2> <read_every> [out.bin]\n";
        return 2;
    }
    const size_t n = std::strtoul(argv[1], nullptr, 10);
    const int steps = std::atoi(argv[2]), prec = std::atoi(argv[3]), every = std::atoi(argv[4]);
    const double frac = 0.665, rho_bar = 2.0, G = 0.9;

    Geometry geom(n, n);
    RNG gen(2024);
    Fluid fluid_1("water");
    Fluid fluid_2("oil");
    fluid_1.initialize_randomly_perturbed_density(geom, frac * rho_bar, gen, true);
    fluid_2.initialize_randomly_perturbed_density(geom, (1 - frac) * rho_bar, gen, true);
    fluid_1.initialize_fluid_repulsion(G);
    fluid_2.initialize_fluid_repulsion(G);
    fluid_1.set_multicomponent_interactions(0.0);
    fluid_2.set_multicomponent_interactions(0.0);
    if (prec != SCL_FP64) {
        fluid_1.set_precision(prec, frac * rho_bar);
        fluid_2.set_precision(prec, (1 - frac) * rho_bar);
    }
    LBM lbm(geom);
    lbm.compute_solid_surface_force(geom, fluid_1, fluid_2);

    auto step = [&]() {
        fluid_1.compute_density();
        fluid_2.compute_density();
        lbm.compute_fluid_repulsive_interactions(geom, fluid_1, fluid_2);
        lbm.compute_equilibrium_velocities(geom, fluid_1, fluid_2);
        lbm.collide(fluid_1, fluid_2);
        lbm.stream(geom, fluid_1, fluid_2);
    };
    double probe = 0.0;
    step();
    sclbm_flush();
    scl_synchronize();
    const auto t0 = std::chrono::steady_clock::now();
    for (int it = 0; it < steps; ++it) {
        step();
        if (every > 0 && (it + 1) % every == 0) {
            const Fluid& a = fluid_1;
            const Fluid& b = fluid_2;
            probe += a.get_f_dist()[0] + b.get_f_dist()[0];
        } else if (every < 0 && (it + 1) % (-every) == 0) {
            sclbm_flush();
        }
    }
    sclbm_flush();
    scl_synchronize();
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("{\"n\": %zu, \"steps\": %d, \"precision\": %d, \"read_every\": %d, \"seconds\": %.6f, "
                "\"mlups\": %.1f, \"probe\": %.17g}\n",
                n, steps, prec, every, s, (double)n * n * steps / s / 1e6, probe);
    if (argc > 5) {
        std::ofstream out(argv[5], std::ios::binary);
        for (const Fluid* fl : {&fluid_1, &fluid_2}) {
            const std::vector<double>& f = fl->get_f_dist();
            out.write(reinterpret_cast<const char*>(f.data()), sizeof(double) * f.size());
        }
    }
    return 0;
}
