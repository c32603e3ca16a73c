/*
 * lbm.h -- drop-in for the reference's LBM operator class (include/lbm.h:23-135
 * of LBMFluids/sc_lbm), executing on an MI355X through libsclbm.so.
 *
 * Every method keeps the reference's signature and meaning.  The operators
 * of one time step are recorded as they are called; when stream() closes a
 * step whose calls form the reference's step loop
 *   one component:  collide(geom, f); [add_volume_force(geom, f, F)]; stream(geom, f)
 *   two components: a.compute_density(); b.compute_density();
 *                   compute_fluid_repulsive_interactions(geom, a, b);
 *                   compute_equilibrium_velocities(geom, a, b); collide(a, b);
 *                   [add_volume_force(...) at most once per component];
 *                   stream(geom, a, b)
 * the whole step runs as one fused kernel; any other sequence runs operator
 * by operator, each as its own kernel with the reference's arithmetic.  Both
 * give the reference's results bit for bit in FP64.  Any host read of a
 * Fluid array first completes the recorded operators.
 */
#ifndef SCLBM_LBM_H
#define SCLBM_LBM_H

#include <algorithm>
#include <iostream>
#include <vector>

#include "common.h"
#include "fluid.h"
#include "geometry.h"
#include "io_operations/lbm_io.h"
#include "rng.h"
#include "sc_lbm_amd.h"
#include "utils.h"

class LBM {
public:
    LBM() = delete;
    /// Stores the lattice size (include/lbm.h:30-37)
    LBM(const Geometry& geom);

    /// Initialisers on the populations (src/lbm.cpp:11-99)
    void initialize_droplet(const Geometry& geom, Fluid& bulk, Fluid& droplet, const double rho_bulk,
                            const double rho_droplet, const double rho_b_in_d, const double rho_d_in_b,
                            const double xc, const double yc, const double radius);
    void initialize_fluid_rectangle(const Geometry& geom, Fluid& bulk, Fluid& droplet, const double rho_bulk,
                                    const double rho_droplet, const double rho_b_in_d, const double rho_d_in_b,
                                    const double xc, const double yc, const double half_Lx, const double half_Ly);

    /// Operators (src/lbm.cpp:102-435)
    void compute_solid_surface_force(const Geometry&, Fluid&, Fluid&);
    void compute_fluid_repulsive_interactions(const Geometry&, Fluid&, Fluid&);
    void compute_equilibrium_velocities(Geometry& geom, Fluid&, Fluid&);
    void collide(const Geometry& geom, Fluid&);
    void collide(Fluid&, Fluid&);
    void add_volume_force(const Geometry&, Fluid&, const std::vector<double>&);
    void add_volume_force(const Geometry&, Fluid&, Fluid&, const std::vector<double>&);
    void stream(const Geometry&, Fluid&);
    void stream(const Geometry&, Fluid&, Fluid&);

private:
    size_t Nx = 0, Ny = 0, Ntot = 0, Ndir = 9;
};

/// Completes every recorded operator (done automatically before any host
/// read; exposed for timing loops).
void sclbm_flush();

#endif
