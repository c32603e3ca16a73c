// fluid.cpp -- Fluid drop-in (see include/fluid.h).
//
// Host mirrors follow the reference's member sizes (src/fluid.cpp:43-151:
// which initialiser sizes which array); the arrays themselves live in the
// device field.  Initialisers write the populations on the host exactly as
// the reference does (f_j = rho/9 on fluid sites, the same loop order and
// RNG calls), and the next device operation uploads them.
#include <cerrno>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <sstream>
#include <stdexcept>

#include "engine.h"
#include "fluid.h"
#include "lbm.h"

Fluid::Fluid(const std::string& s, const double speed, const double rel) : name(s), cs2(speed), tau(rel) {
    omega = 1 / tau;              // include/fluid.h:62
    nu = cs2 * (tau - 0.5);
    for (int a = 0; a < A_COUNT; ++a) {
        host_ok[a] = true;
        dev_ok[a] = true;
    }
}

Fluid::Fluid(const Fluid& other)
    : name(other.name), cs2(other.cs2), tau(other.tau), nu(other.nu), omega(other.omega), Gsolid(other.Gsolid),
      Gfluid_repulsion(other.Gfluid_repulsion), Nx(other.Nx), Ny(other.Ny), Ntot(other.Ntot),
      precision(other.precision), rho0(other.rho0), solid_force_from_mask(other.solid_force_from_mask),
      solid_force_set(other.solid_force_set) {
    for (int a = 0; a < A_COUNT; ++a) {
        other.sync_host(a);
        arr[a] = other.arr[a];
        host_ok[a] = true;
        dev_ok[a] = false;
    }
    dev_geom = other.dev_geom;
}

Fluid& Fluid::operator=(const Fluid& other) {
    if (this == &other) return *this;
    for (int a = 0; a < A_COUNT; ++a) other.sync_host(a);
    release_device();
    name = other.name;
    cs2 = other.cs2;
    tau = other.tau;
    nu = other.nu;
    omega = other.omega;
    Gsolid = other.Gsolid;
    Gfluid_repulsion = other.Gfluid_repulsion;
    Nx = other.Nx;
    Ny = other.Ny;
    Ntot = other.Ntot;
    precision = other.precision;
    rho0 = other.rho0;
    solid_force_from_mask = other.solid_force_from_mask;
    solid_force_set = other.solid_force_set;
    for (int a = 0; a < A_COUNT; ++a) {
        arr[a] = other.arr[a];
        host_ok[a] = true;
        dev_ok[a] = false;
    }
    dev_geom = other.dev_geom;
    return *this;
}

Fluid::~Fluid() { release_device(); }

void Fluid::release_device() {
    sclbm_detail::forget(this);
    if (dev) {
        scl_field_destroy(dev);
        dev = nullptr;
    }
}

// ---- mirrors --------------------------------------------------------------

void Fluid::sync_host(int a) const {
    if (host_ok[a]) return;
    sclbm_detail::flush_log();
    if (host_ok[a]) return;
    std::vector<double>& v = arr[a];
    const size_t n = (a == A_F || a == A_FEQ) ? Ndir * Ntot : Ntot;
    // the deviation forms keep populations and u_eq only
    const bool kept = precision == SCL_FP64 || a == A_F || a == A_UEQX || a == A_UEQY;
    if (v.size() >= n && n > 0 && dev && kept) {
        // the reference appends to F_solid; the device holds the first Ntot
        std::vector<double> tmp(n);
        sclbm_detail::raise_status(scl_field_download(dev, a, tmp.data(), n));
        std::copy(tmp.begin(), tmp.end(), v.begin());
    }
    host_ok[a] = true;
}

std::vector<double>& Fluid::host(int a, bool hand_out) {
    sclbm_detail::flush_log();
    sync_host(a);
    if (hand_out) {
        dev_ok[a] = false;
        if (a == A_FSX || a == A_FSY) solid_force_from_mask = false;
    }
    return arr[a];
}

const std::vector<double>& Fluid::chost(int a) const {
    sclbm_detail::flush_log();
    sync_host(a);
    return arr[a];
}

scl_field* Fluid::device_field(const Geometry& geom) {
    if (dev) return dev;
    scl_field* f = nullptr;
    sclbm_detail::raise_status(scl_field_create(geom.device(), precision, tau, rho0, &f));
    dev = f;
    dev_geom = &geom;
    sclbm_detail::raise_status(scl_field_set_interactions(dev, Gsolid, Gfluid_repulsion));
    for (int a = 0; a < A_COUNT; ++a) dev_ok[a] = false;
    return dev;
}

// The initialisers only bind the geometry: the device field is created by the
// first device operation, so building and initialising a Fluid (and reading
// its host arrays) needs no GPU.
void Fluid::bind_geometry(const Geometry& geom) {
    if (!dev) dev_geom = &geom;
}

// Upload every array whose device copy is stale (host copies that were
// written, or handed out, since the last device operation).
void Fluid::push_device() {
    if (!dev) throw std::logic_error("Fluid " + name + " is not initialised (no *_ini call)");
    sclbm_detail::raise_status(scl_field_set_interactions(dev, Gsolid, Gfluid_repulsion));
    for (int a = 0; a < A_COUNT; ++a) {
        if (dev_ok[a]) continue;
        const size_t n = (a == A_F || a == A_FEQ) ? Ndir * Ntot : Ntot;
        if (arr[a].size() >= n && n > 0) {
            if (precision != SCL_FP64 && a != A_F && a != A_UEQX && a != A_UEQY) {
                dev_ok[a] = true;   // deviation forms keep only populations and u_eq
                continue;
            }
            sclbm_detail::raise_status(scl_field_upload(dev, a, arr[a].data(), n));
        }
        dev_ok[a] = true;
    }
}

void Fluid::device_wrote() const {
    for (int a = 0; a < A_COUNT; ++a) host_ok[a] = false;
}

void Fluid::set_precision(int prec, double r0) {
    if (prec != SCL_FP64 && prec != SCL_FP32_DEV && prec != SCL_FP16_DEV)
        throw std::invalid_argument("unknown precision");
    sclbm_detail::flush_log();
    for (int a = 0; a < A_COUNT; ++a) sync_host(a);
    const Geometry* g = dev_geom;
    if (dev) {
        scl_field_destroy(dev);
        dev = nullptr;
    }
    precision = prec;
    rho0 = r0;
    for (int a = 0; a < A_COUNT; ++a) dev_ok[a] = false;
    dev_geom = g;   // the device field is created again by the next device operation
}

// ---- initialisers (src/fluid.cpp:43-151) ----------------------------------

void Fluid::zero_density_ini(const Geometry& geom) {
    Nx = geom.Nx();
    Ny = geom.Ny();
    Ntot = Nx * Ny;
    for (int a : {A_RHO, A_UX, A_UY, A_UEQX, A_UEQY, A_FRX, A_FRY}) {
        sync_host(a);
        arr[a].resize(Ntot, 0.0);
        dev_ok[a] = false;
    }
    for (int a : {A_F, A_FEQ}) {
        sync_host(a);
        arr[a].resize(Ntot * Ndir, 0.0);
        dev_ok[a] = false;
    }
    bind_geometry(geom);
}

void Fluid::simple_ini(const Geometry& geom, const double rho_0, const bool mcmp) {
    Nx = geom.Nx();
    Ny = geom.Ny();
    Ntot = Nx * Ny;
    const double rho_factor = static_cast<double>(Ndir);
    for (int a : {A_RHO, A_UX, A_UY}) {
        sync_host(a);
        arr[a].resize(Ntot, 0.0);
        dev_ok[a] = false;
    }
    if (mcmp)
        for (int a : {A_UEQX, A_UEQY, A_FRX, A_FRY}) {
            sync_host(a);
            arr[a].resize(Ntot, 0.0);
            dev_ok[a] = false;
        }
    sync_host(A_F);
    std::vector<double>& f = arr[A_F];
    f.resize(Ntot * Ndir, 0.0);
    for (size_t i = 0; i < Nx; ++i)
        for (size_t j = 0; j < Ny; ++j)
            if (geom(i, j) == 1)
                for (size_t k = 0; k < Ndir; ++k) f.at(k * Nx * Ny + j * Nx + i) = rho_0 / rho_factor;
    dev_ok[A_F] = false;
    sync_host(A_FEQ);
    arr[A_FEQ].resize(Ntot * Ndir, 0.0);
    dev_ok[A_FEQ] = false;
    bind_geometry(geom);
}

void Fluid::initialize_randomly_perturbed_density(const Geometry& geom, const double rho_0, RNG& gen,
                                                  const bool mcmp) {
    Nx = geom.Nx();
    Ny = geom.Ny();
    Ntot = Nx * Ny;
    const double rho_factor = static_cast<double>(Ndir);
    const double rand_max = rho_0 / 1000.0;
    for (int a : {A_RHO, A_UX, A_UY}) {
        sync_host(a);
        arr[a].resize(Ntot, 0.0);
        dev_ok[a] = false;
    }
    if (mcmp)
        for (int a : {A_UEQX, A_UEQY, A_FRX, A_FRY}) {
            sync_host(a);
            arr[a].resize(Ntot, 0.0);
            dev_ok[a] = false;
        }
    sync_host(A_F);
    std::vector<double>& f = arr[A_F];
    f.resize(Ntot * Ndir, 0.0);
    for (size_t i = 0; i < Nx; ++i)
        for (size_t j = 0; j < Ny; ++j)
            if (geom(i, j) == 1) {
                const double rho_rand = rho_0 + gen.get_random(0.0, rand_max);
                for (size_t k = 0; k < Ndir; ++k) f.at(k * Nx * Ny + j * Nx + i) = rho_rand / rho_factor;
            }
    dev_ok[A_F] = false;
    sync_host(A_FEQ);
    arr[A_FEQ].resize(Ntot * Ndir, 0.0);
    dev_ok[A_FEQ] = false;
    bind_geometry(geom);
}

void Fluid::set_multicomponent_interactions(const double Gs) {
    std::cout << "This function is depracted and will be removed in future releases" << std::endl;
    Gsolid = Gs;
}

void Fluid::initialize_fluid_repulsion(const double Gf) {
    Gfluid_repulsion = Gf;
    for (int a : {A_FRX, A_FRY}) {
        sync_host(a);
        arr[a].resize(Nx * Ny, 0.0);
        dev_ok[a] = false;
    }
}

// src/fluid.cpp:154-160: append (-1.0*Gs)*F to F_solid.  Values supplied by
// the caller are not the mask-derived ones the fused kernel recomputes, so
// this component then runs operator by operator.
void Fluid::add_surface_forces(const std::vector<double> Fxs, const std::vector<double> Fys) {
    sync_host(A_FSX);
    sync_host(A_FSY);
    const bool first = arr[A_FSX].empty();
    for (double el : Fxs) arr[A_FSX].push_back(-1.0 * Gsolid * el);
    for (double el : Fys) arr[A_FSY].push_back(-1.0 * Gsolid * el);
    if (first) {
        dev_ok[A_FSX] = false;
        dev_ok[A_FSY] = false;
        solid_force_from_mask = false;
    }
    solid_force_set = true;
}

// ---- macroscopic operators: recorded, executed on the device ---------------

void Fluid::compute_density() { sclbm_detail::record_fluid_op(sclbm_detail::OP_DENSITY, this, nullptr); }

void Fluid::compute_velocities(const Geometry& geom) {
    sclbm_detail::record_fluid_op(sclbm_detail::OP_VELOCITIES, this, &geom);
}

void Fluid::compute_macroscopic(const Geometry& geom) {
    compute_density();
    compute_velocities(geom);
}

void Fluid::compute_f_equilibrium(const Geometry& geom) {
    sclbm_detail::record_fluid_op(sclbm_detail::OP_FEQ_GEOM, this, &geom);
}

void Fluid::compute_f_equilibrium() { sclbm_detail::record_fluid_op(sclbm_detail::OP_FEQ_MC, this, nullptr); }

// ---- text I/O (src/fluid.cpp:285-363; 6 significant digits) ---------------

void Fluid::restart_from_file(const std::string& fname) {
    sync_host(A_F);
    std::vector<double>& f = arr[A_F];
    for (size_t id = 0; id < Ndir; ++id) {
        const std::string fn = fname + "_" + std::to_string(id) + ".txt";
        std::ifstream in(fn);
        if (!in) {
            std::cerr << "Error opening file " << fn << std::endl;
            throw std::ios_base::failure(std::strerror(errno));
        }
        // every line of the file is one row y (values separated by spaces);
        // a missing row is std::out_of_range, as the reference's .at()
        std::vector<std::vector<double>> rows;
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ss(line);
            std::vector<double> r;
            double v;
            while (ss >> v) r.push_back(v);
            rows.push_back(std::move(r));
        }
        for (size_t yj = 0; yj < Ny; ++yj) {
            const std::vector<double>& r = rows.at(yj);
            for (size_t xi = 0; xi < r.size(); ++xi) f.at(id * Ntot + yj * Nx + xi) = r[xi];
        }
    }
    dev_ok[A_F] = false;
}

void Fluid::write_var(const std::vector<double>& variable, const std::string& fname) const {
    std::ofstream out(fname, std::ios_base::out | std::ios_base::trunc);
    if (!out) {
        std::cerr << "Error opening file " << fname << std::endl;
        throw std::ios_base::failure(std::strerror(errno));
    }
    for (size_t y = 0; y < Ny; ++y) {
        std::copy(variable.begin() + y * Nx, variable.begin() + (y + 1) * Nx,
                  std::ostream_iterator<double>(out, " "));
        out << '\n';
    }
}

void Fluid::write_var(const std::vector<double>& variable, const std::string& fname, const bool) const {
    std::vector<double> plane(Nx * Ny);
    for (size_t i = 0; i < Ndir; ++i) {
        std::copy(variable.begin() + i * Nx * Ny, variable.begin() + (i + 1) * Nx * Ny, plane.begin());
        write_var(plane, fname + "_" + std::to_string(i) + ".txt");
    }
}

void Fluid::save_state(const std::string& frho, const std::string& fux, const std::string& fuy, const int step,
                       const Geometry& geom) {
    compute_density();
    write_var(get_rho(), frho + "_" + std::to_string(step) + ".txt");
    compute_velocities(geom);
    write_var(get_ux(), fux + "_" + std::to_string(step) + ".txt");
    write_var(get_uy(), fuy + "_" + std::to_string(step) + ".txt");
}
