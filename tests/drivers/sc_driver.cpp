// sc_driver.cpp -- a driver written against the drop-in C++ API
// (include/lbm.h, fluid.h, geometry.h) the way the reference's own drivers
// are (benchmarks/single_phase_one_component_flows/laminar_flow_x_dir.cpp,
// benchmarks/two_phase_two_components_flows/phase_separation/phase_separation.cpp,
// tests/lbm/lbm_tests.cpp).  It reads a case file written by
// tests/drivers/case_io.py, runs the step loop and writes every array of
// every component as raw doubles, for the parity tests.
//
// usage: sc_driver <case.bin> <out.bin>
// modes: 0 the reference loop (runs fused), 1 the same loop with a host read
//        of the density after compute_density (runs operator by operator),
//        2 the reference loop with a host round trip of the populations
//        every 7 steps, 3 the two-component loop without
//        compute_solid_surface_force (the reference throws
//        std::out_of_range; exit status 3).
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "lbm.h"

namespace {

template <class T>
void rd(std::ifstream& in, T* p, size_t n) {
    in.read(reinterpret_cast<char*>(p), sizeof(T) * n);
    if (!in) throw std::runtime_error("short case file");
}

void wr(std::ofstream& out, const std::vector<double>& v) {
    const uint64_t n = v.size();
    out.write(reinterpret_cast<const char*>(&n), sizeof n);
    out.write(reinterpret_cast<const char*>(v.data()), sizeof(double) * v.size());
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: sc_driver <case.bin> <out.bin>\n";
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    int32_t hdr[8];
    rd(in, hdr, 8);
    if (hdr[0] != 0x424c4353) throw std::runtime_error("bad case file");
    const int nx = hdr[1], ny = hdr[2], ncomp = hdr[3], steps = hdr[4], mode = hdr[5];
    const bool has_force[2] = {hdr[6] != 0, hdr[7] != 0};
    double tau[2], gf[2], gs[2], force[2][9];
    rd(in, tau, 2);
    rd(in, gf, 2);
    rd(in, gs, 2);
    rd(in, &force[0][0], 18);
    std::vector<int32_t> mask((size_t)nx * ny);
    rd(in, mask.data(), mask.size());
    const size_t n9 = (size_t)9 * nx * ny;
    std::vector<std::vector<double>> f0(ncomp, std::vector<double>(n9));
    for (int c = 0; c < ncomp; ++c) rd(in, f0[c].data(), n9);

    Geometry geom(nx, ny);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
            if (mask[(size_t)y * nx + x] == 0) geom.set_node_solid(x, y);

    std::vector<double> fa(force[0], force[0] + 9), fb(force[1], force[1] + 9);
    std::ofstream out(argv[2], std::ios::binary);

    if (ncomp == 1) {
        Fluid fluid("fluid", 1. / 3, tau[0]);
        fluid.simple_ini(geom, 0.0);
        fluid.get_f_dist() = f0[0];
        LBM lbm(geom);
        for (int it = 0; it < steps; ++it) {
            lbm.collide(geom, fluid);
            if (mode == 1) (void)fluid.get_f_eq_dist()[0];
            if (has_force[0]) lbm.add_volume_force(geom, fluid, fa);
            lbm.stream(geom, fluid);
            if (mode == 2 && it % 7 == 3) {
                std::vector<double>& f = fluid.get_f_dist();
                f[0] = f[0] + 0.0;
            }
        }
        wr(out, fluid.get_f_dist());
        wr(out, fluid.get_rho());
        wr(out, fluid.get_ux());
        wr(out, fluid.get_uy());
        wr(out, fluid.get_f_eq_dist());
        return 0;
    }

    Fluid A("fluid_1", 1. / 3, tau[0]);
    Fluid B("fluid_2", 1. / 3, tau[1]);
    A.zero_density_ini(geom);
    B.zero_density_ini(geom);
    A.initialize_interactions(gs[0], gf[0]);
    B.initialize_interactions(gs[1], gf[1]);
    A.get_f_dist() = f0[0];
    B.get_f_dist() = f0[1];
    LBM lbm(geom);
    if (mode == 3) {
        try {
            A.compute_density();
            B.compute_density();
            lbm.compute_fluid_repulsive_interactions(geom, A, B);
            lbm.compute_equilibrium_velocities(geom, A, B);
        } catch (const std::out_of_range& e) {
            std::cout << "out_of_range: " << e.what() << std::endl;
            return 3;
        }
        return 0;
    }
    lbm.compute_solid_surface_force(geom, A, B);
    for (int it = 0; it < steps; ++it) {
        A.compute_density();
        B.compute_density();
        if (mode == 1) (void)A.get_rho()[0];
        lbm.compute_fluid_repulsive_interactions(geom, A, B);
        lbm.compute_equilibrium_velocities(geom, A, B);
        lbm.collide(A, B);
        if (has_force[0] && has_force[1] && fa == fb) {
            lbm.add_volume_force(geom, A, B, fa);
        } else {
            if (has_force[0]) lbm.add_volume_force(geom, A, fa);
            if (has_force[1]) lbm.add_volume_force(geom, B, fb);
        }
        lbm.stream(geom, A, B);
        if (mode == 2 && it % 7 == 3) {
            std::vector<double>& f = B.get_f_dist();
            f[0] = f[0] + 0.0;
        }
    }
    for (Fluid* fl : {&A, &B}) {
        wr(out, fl->get_f_dist());
        wr(out, fl->get_rho());
        wr(out, fl->get_ux());
        wr(out, fl->get_uy());
        wr(out, fl->get_f_eq_dist());
        wr(out, fl->get_u_eq_x());
        wr(out, fl->get_u_eq_y());
        wr(out, fl->get_repulsive_force_x());
        wr(out, fl->get_repulsive_force_y());
        wr(out, fl->get_fluid_solid_force_x());
        wr(out, fl->get_fluid_solid_force_y());
    }
    return 0;
}
