/*
 * fluid.h -- drop-in for the reference's Fluid (include/fluid.h:52-320 of
 * LBMFluids/sc_lbm): one fluid component with its populations and the
 * per-site arrays of the step, stored flat as in the reference
 * (site y*Nx + x, population i at i*Nx*Ny + site).
 *
 * The arrays live on the GPU.  Each getter returns a host mirror that is
 * brought up to date first; a non-const getter also hands the array to the
 * caller, so the next device operation uploads it again.  A reference
 * obtained from a getter is the mirror at that moment: writes through it are
 * seen by the device if they happen before the next LBM / Fluid operation.
 */
#ifndef SCLBM_FLUID_H
#define SCLBM_FLUID_H

#include <algorithm>
#include <cstddef>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "common.h"
#include "geometry.h"
#include "io_operations/lbm_io.h"
#include "rng.h"
#include "sc_lbm_amd.h"
#include "utils.h"

class LBM;
namespace sclbm_detail {
struct Engine;
}

class Fluid {
public:
    /// Name, squared lattice speed of sound, relaxation time; omega = 1/tau
    /// (include/fluid.h:60-62)
    Fluid(const std::string& s, const double speed, const double rel);
    explicit Fluid(const std::string& s) : Fluid(s, 1. / 3, 1.0) {}
    Fluid() : Fluid("fluid", 1. / 3, 1.0) {}
    Fluid(const Fluid& other);
    Fluid& operator=(const Fluid& other);
    ~Fluid();

    /// Initialisers (src/fluid.cpp:43-151)
    void zero_density_ini(const Geometry& geom);
    void simple_ini(const Geometry& geom, const double rho_0, const bool mcmp = false);
    void initialize_randomly_perturbed_density(const Geometry& geom, const double rho_0, RNG& gen,
                                               const bool mcmp = false);

    /// Interaction parameters (include/fluid.h:97-117)
    void set_multicomponent_interactions(const double Gs);
    void initialize_interactions(const double Gs, const double Gf) {
        Gsolid = Gs;
        initialize_fluid_repulsion(Gf);
    }
    void add_surface_forces(const std::vector<double> Fxs, const std::vector<double> Fys);
    void initialize_fluid_repulsion(const double Gf);

    /// Macroscopic quantities (src/fluid.cpp:175-207)
    void compute_density();
    void compute_velocities(const Geometry& geom);
    void compute_macroscopic(const Geometry& geom);

    /// Equilibrium distributions (src/fluid.cpp:210-278)
    void compute_f_equilibrium(const Geometry& geom);
    void compute_f_equilibrium();

    /// Getters (include/fluid.h:150-201)
    std::vector<double>& get_f_dist() { return host(A_F, true); }
    std::vector<double>& get_f_eq_dist() { return host(A_FEQ, true); }
    std::vector<double>& get_fluid_solid_force_x() { return host(A_FSX, true); }
    std::vector<double>& get_fluid_solid_force_y() { return host(A_FSY, true); }
    std::vector<double>& get_repulsive_force_x() { return host(A_FRX, true); }
    std::vector<double>& get_repulsive_force_y() { return host(A_FRY, true); }
    std::vector<double>& get_rho() { return host(A_RHO, true); }
    std::vector<double>& get_ux() { return host(A_UX, true); }
    std::vector<double>& get_uy() { return host(A_UY, true); }
    std::vector<double>& get_u_eq_x() { return host(A_UEQX, true); }
    std::vector<double>& get_u_eq_y() { return host(A_UEQY, true); }
    const std::vector<double>& get_f_dist() const { return chost(A_F); }
    const std::vector<double>& get_f_eq_dist() const { return chost(A_FEQ); }
    const std::vector<double>& get_fluid_solid_force_x() const { return chost(A_FSX); }
    const std::vector<double>& get_fluid_solid_force_y() const { return chost(A_FSY); }
    const std::vector<double>& get_rho() const { return chost(A_RHO); }
    const std::vector<double>& get_ux() const { return chost(A_UX); }
    const std::vector<double>& get_uy() const { return chost(A_UY); }
    const std::vector<double>& get_u_eq_x() const { return chost(A_UEQX); }
    const std::vector<double>& get_u_eq_y() const { return chost(A_UEQY); }

    size_t get_Nx() const { return Nx; }
    size_t get_Ny() const { return Ny; }
    double get_omega() const { return omega; }
    double get_repulsive_g_fluid() const { return Gfluid_repulsion; }
    std::vector<double> get_wrts() const { return {wrt0, wrt1, wrt2}; }

    /// Restart from the text files written by write_f (src/fluid.cpp:285-300)
    void restart_from_file(const std::string& fname);

    /// Text output, 6 significant digits (src/fluid.cpp:303-363)
    void write_density(const std::string& fname) {
        compute_density();
        write_var(get_rho(), fname);
    }
    void write_ux(const std::string& fname, const Geometry& geom) {
        compute_velocities(geom);
        write_var(get_ux(), fname);
    }
    void write_uy(const std::string& fname, const Geometry& geom) {
        compute_velocities(geom);
        write_var(get_uy(), fname);
    }
    void write_F_repulsive_x(const std::string& fname) { write_var(chost(A_FRX), fname); }
    void write_F_repulsive_y(const std::string& fname) { write_var(chost(A_FRY), fname); }
    void write_u_eq_x(const std::string& fname) { write_var(chost(A_UEQX), fname); }
    void write_u_eq_y(const std::string& fname) { write_var(chost(A_UEQY), fname); }
    void write_feq(const std::string& fname) const { write_var(chost(A_FEQ), fname, true); }
    void write_f(const std::string& fname) const { write_var(chost(A_F), fname, true); }
    void save_state(const std::string& frho, const std::string& fux, const std::string& fuy, const int step,
                    const Geometry& geom);

    /// Extension (no reference counterpart): store the populations of this
    /// component as SCL_FP32_DEV or SCL_FP16_DEV deviations from w_i*rho0.
    /// Such a component runs the canonical step sequences only (fused).
    void set_precision(int precision, double rho0 = 1.0);

private:
    friend class LBM;
    friend struct sclbm_detail::Engine;
    enum Arr { A_F = 0, A_FEQ, A_RHO, A_UX, A_UY, A_UEQX, A_UEQY, A_FRX, A_FRY, A_FSX, A_FSY, A_COUNT };

    std::string name;
    const size_t Ndir = 9;
    double cs2 = 0.0, tau = 0.0, nu = 0.0, omega = 0.0;
    double Gsolid = 0.0, Gfluid_repulsion = 0.0;
    double wrt0 = 4.0 / 9.0, wrt1 = 1.0 / 9.0, wrt2 = 1.0 / 36.0;
    size_t Nx = 0, Ny = 0, Ntot = 0;
    int precision = SCL_FP64;
    double rho0 = 1.0;

    // host mirrors, sized exactly as the reference sizes its members
    mutable std::vector<double> arr[A_COUNT];
    // validity of each array on the host and on the device
    mutable bool host_ok[A_COUNT];
    mutable bool dev_ok[A_COUNT];
    // fluid-solid force written by LBM::compute_solid_surface_force (true)
    // or through the host (false: the fused kernel, which recomputes it from
    // the mask, must not be used)
    bool solid_force_from_mask = true;
    bool solid_force_set = false;

    // device state
    mutable scl_field* dev = nullptr;
    mutable const Geometry* dev_geom = nullptr;

    std::vector<double>& host(int a, bool hand_out);
    const std::vector<double>& chost(int a) const;
    void sync_host(int a) const;
    scl_field* device_field(const Geometry& geom);
    void bind_geometry(const Geometry& geom);
    void push_device();
    void device_wrote() const;
    void release_device();
    void write_var(const std::vector<double>& variable, const std::string& fname) const;
    void write_var(const std::vector<double>& variable, const std::string& fname, const bool is_3D) const;
};

// the reference's fluid.h includes lbm.h (include/fluid.h:9): drivers that
// include only fluid.h use LBM too
#include "lbm.h"

#endif
