// dropin_host.cpp -- drives the drop-in classes (include/geometry.h, lbm.h,
// fluid.h) through the calls a reference driver makes before and around its
// step loop: the Geometry builders, the LBM / Fluid initialisers, the text
// output and restart, and optionally the two-component step loop.  Used by
// the parity tests against the reference's golden vectors; everything up to
// the step loop runs on the host (no GPU).
//
// usage: dropin_host <out.bin> <command>...
//   G nx ny                       Geometry(nx, ny)
//   W dh x|y                      add_walls(dh, "x" / "y")
//   C D xc yc                     add_circle
//   R Lx Ly xc yc                 add_rectangle
//   Q L xc yc                     add_square
//   E Lx Ly xc yc                 add_ellipse
//   AD type n p1..pn x0 x1 y0 y1 dx dy    add_array by gaps
//   AN type n p1..pn x0 x1 y0 y1 nx ny    add_array by numbers of objects
//   F tauA tauB                   two fluids, zero_density_ini
//   ID rb rd rbd rdb xc yc r      LBM::initialize_droplet(geom, A, B, ...)
//   IR rb rd rbd rdb xc yc hx hy  LBM::initialize_fluid_rectangle(geom, A, B, ...)
//   IS rhoA rhoB                  simple_ini(geom, rho, true) on A and B
//   IP rhoA rhoB seed             initialize_randomly_perturbed_density on A
//                                 then B, one seeded RNG
//   S GsA GsB GfA GfB steps fa[9] fb[9]   the reference's two-component loop
//                                 (a zero force vector is not applied)
//   T prefix                      A.write_f / B.write_f to prefix_A / prefix_B,
//                                 then two new fluids restart_from_file them
//                                 and replace A and B
// out.bin: int32 status (0 ok, 1 runtime_error, 2 invalid_argument,
// 3 out_of_range, 4 ios_base::failure, 5 other exception), int32 nx, ny,
// the mask (nx*ny int32), int32 number of fluids, then 9*nx*ny doubles per
// fluid (the populations).  A failing command ends the run.
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm.h"

namespace {

struct Args {
    int argc;
    char** argv;
    int i;
    std::string s() {
        if (i >= argc) throw std::logic_error("missing argument");
        return argv[i++];
    }
    size_t z() { return (size_t)std::stoull(s()); }
    double d() { return std::strtod(s().c_str(), nullptr); }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: dropin_host <out.bin> <command>...\n";
        return 2;
    }
    Args a{argc, argv, 2};
    std::unique_ptr<Geometry> geom;
    std::unique_ptr<Fluid> A, B;
    int status = 0;
    try {
        while (a.i < argc) {
            const std::string c = a.s();
            if (c == "G") {
                const size_t nx = a.z(), ny = a.z();
                geom.reset(new Geometry(nx, ny));
            } else if (c == "W") {
                const size_t dh = a.z();
                geom->add_walls(dh, a.s());
            } else if (c == "C") {
                const size_t D = a.z(), xc = a.z(), yc = a.z();
                geom->add_circle(D, xc, yc);
            } else if (c == "R" || c == "E") {
                const size_t lx = a.z(), ly = a.z(), xc = a.z(), yc = a.z();
                if (c == "R") geom->add_rectangle(lx, ly, xc, yc);
                else geom->add_ellipse(lx, ly, xc, yc);
            } else if (c == "Q") {
                const size_t L = a.z(), xc = a.z(), yc = a.z();
                geom->add_square(L, xc, yc);
            } else if (c == "AD" || c == "AN") {
                const std::string type = a.s();
                std::vector<size_t> p(a.z());
                for (size_t& v : p) v = a.z();
                std::vector<std::vector<size_t>> bounds(2, std::vector<size_t>(2));
                bounds[0][0] = a.z();
                bounds[0][1] = a.z();
                bounds[1][0] = a.z();
                bounds[1][1] = a.z();
                const size_t u = a.z(), v = a.z();
                if (c == "AD") geom->add_array(p, bounds, u, v, type);
                else geom->add_array(p, bounds, std::vector<size_t>{u, v}, type);
            } else if (c == "F") {
                const double ta = a.d(), tb = a.d();
                A.reset(new Fluid("fluid_A", 1. / 3, ta));
                B.reset(new Fluid("fluid_B", 1. / 3, tb));
                A->zero_density_ini(*geom);
                B->zero_density_ini(*geom);
            } else if (c == "ID" || c == "IR") {
                double v[8];
                const int n = c == "ID" ? 7 : 8;
                for (int k = 0; k < n; ++k) v[k] = a.d();
                LBM lbm(*geom);
                if (c == "ID") lbm.initialize_droplet(*geom, *A, *B, v[0], v[1], v[2], v[3], v[4], v[5], v[6]);
                else lbm.initialize_fluid_rectangle(*geom, *A, *B, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
            } else if (c == "IS") {
                const double ra = a.d(), rb = a.d();
                A->simple_ini(*geom, ra, true);
                B->simple_ini(*geom, rb, true);
            } else if (c == "IP") {
                const double ra = a.d(), rb = a.d();
                RNG gen((std::uint32_t)a.z());
                A->initialize_randomly_perturbed_density(*geom, ra, gen, true);
                B->initialize_randomly_perturbed_density(*geom, rb, gen, true);
            } else if (c == "S") {
                const double gsa = a.d(), gsb = a.d(), gfa = a.d(), gfb = a.d();
                const long steps = std::stol(a.s());
                std::vector<double> fa(9), fb(9);
                for (double& v : fa) v = a.d();
                for (double& v : fb) v = a.d();
                auto nonzero = [](const std::vector<double>& f) {
                    for (double v : f)
                        if (v != 0.0) return true;
                    return false;
                };
                A->initialize_interactions(gsa, gfa);
                B->initialize_interactions(gsb, gfb);
                LBM lbm(*geom);
                lbm.compute_solid_surface_force(*geom, *A, *B);
                for (long it = 0; it < steps; ++it) {
                    A->compute_density();
                    B->compute_density();
                    lbm.compute_fluid_repulsive_interactions(*geom, *A, *B);
                    lbm.compute_equilibrium_velocities(*geom, *A, *B);
                    lbm.collide(*A, *B);
                    if (nonzero(fa)) lbm.add_volume_force(*geom, *A, fa);
                    if (nonzero(fb)) lbm.add_volume_force(*geom, *B, fb);
                    lbm.stream(*geom, *A, *B);
                }
            } else if (c == "T") {
                const std::string prefix = a.s();
                A->write_f(prefix + "_A");
                B->write_f(prefix + "_B");
                std::unique_ptr<Fluid> a2(new Fluid("fluid_A", 1. / 3, 1.0)), b2(new Fluid("fluid_B", 1. / 3, 1.0));
                a2->simple_ini(*geom, 0.0);
                b2->simple_ini(*geom, 0.0);
                a2->restart_from_file(prefix + "_A");
                b2->restart_from_file(prefix + "_B");
                A.swap(a2);
                B.swap(b2);
            } else {
                throw std::logic_error("unknown command " + c);
            }
        }
    } catch (const std::invalid_argument& e) {
        status = 2;
        std::cout << "invalid_argument: " << e.what() << std::endl;
    } catch (const std::out_of_range& e) {
        status = 3;
        std::cout << "out_of_range: " << e.what() << std::endl;
    } catch (const std::ios_base::failure& e) {
        status = 4;
        std::cout << "ios_base::failure: " << e.what() << std::endl;
    } catch (const std::runtime_error& e) {
        status = 1;
        std::cout << "runtime_error: " << e.what() << std::endl;
    } catch (const std::exception& e) {
        status = 5;
        std::cout << "exception: " << e.what() << std::endl;
    }
    std::ofstream out(argv[1], std::ios::binary);
    auto put = [&](int32_t v) { out.write(reinterpret_cast<const char*>(&v), sizeof v); };
    put(status);
    const int32_t nx = geom ? (int32_t)geom->Nx() : 0, ny = geom ? (int32_t)geom->Ny() : 0;
    put(nx);
    put(ny);
    if (geom)
        for (int v : geom->get_geom()) put(v);
    const bool have = A && B && status == 0;
    put(have ? 2 : 0);
    if (have)
        for (Fluid* f : {A.get(), B.get()}) {
            const std::vector<double>& v = static_cast<const Fluid&>(*f).get_f_dist();
            out.write(reinterpret_cast<const char*>(v.data()), sizeof(double) * v.size());
        }
    return 0;
}
