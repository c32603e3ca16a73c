/*
 * ref_driver.cpp -- C entry points that drive the compiled reference
 * (LBMFluids/sc_lbm, built from its own sources by oracle/Makefile into
 * oracle/_ref/) through its public C++ API only: Geometry, Fluid, LBM.
 *
 * ORACLE / TEST INFRASTRUCTURE.  Used to write the golden vectors under
 * tests/golden and, optionally, as the CPU baseline.  Nothing in the product
 * links it.  The call sequences are those of the reference's own drivers:
 *   two components -- tests/lbm/lbm_tests.cpp:77-127 and
 *     benchmarks/two_phase_two_components_flows/phase_separation/phase_separation.cpp:99-108
 *   one component  -- benchmarks/single_phase_one_component_flows/laminar_flow_x_dir.cpp:64-68
 */
#include "lbm.h"

#include <memory>
#include <stdexcept>
#include <cstring>

namespace {
struct RefSys {
    int ncomp = 0;
    std::unique_ptr<Geometry> geom;
    std::unique_ptr<Fluid> fl[2];
    std::unique_ptr<LBM> lbm;
};

std::vector<double>* pick(Fluid& f, int field) {
    switch (field) {
        case 0: return &f.get_f_dist();
        case 1: return &f.get_f_eq_dist();
        case 2: return &f.get_rho();
        case 3: return &f.get_ux();
        case 4: return &f.get_uy();
        case 5: return &f.get_u_eq_x();
        case 6: return &f.get_u_eq_y();
        case 7: return &f.get_repulsive_force_x();
        case 8: return &f.get_repulsive_force_y();
        case 9: return &f.get_fluid_solid_force_x();
        case 10: return &f.get_fluid_solid_force_y();
        default: return nullptr;
    }
}
}  // namespace

extern "C" {

void* ref_create(int nx, int ny, int ncomp, const int* mask, const double* tau,
                 const double* g_fluid, const double* g_solid) {
    try {
        RefSys* s = new RefSys;
        s->ncomp = ncomp;
        s->geom.reset(new Geometry((size_t)nx, (size_t)ny));
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                if (mask[(size_t)y * nx + x] == 0) s->geom->set_node_solid(x, y);
        for (int k = 0; k < ncomp; ++k) {
            s->fl[k].reset(new Fluid(k == 0 ? "fluid_1" : "fluid_2", 1. / 3, tau[k]));
            s->fl[k]->zero_density_ini(*s->geom);
            if (ncomp == 2)
                s->fl[k]->initialize_interactions(g_solid ? g_solid[k] : 0.0,
                                                  g_fluid ? g_fluid[k] : 0.0);
        }
        s->lbm.reset(new LBM(*s->geom));
        return s;
    } catch (...) {
        return nullptr;
    }
}

void ref_destroy(void* h) { delete static_cast<RefSys*>(h); }

int ref_set(void* h, int comp, int field, const double* src, size_t n) {
    RefSys* s = static_cast<RefSys*>(h);
    std::vector<double>* v = pick(*s->fl[comp], field);
    if (!v || v->size() != n) return 1;
    std::memcpy(v->data(), src, n * sizeof(double));
    return 0;
}

int ref_get(void* h, int comp, int field, double* dst, size_t n) {
    RefSys* s = static_cast<RefSys*>(h);
    std::vector<double>* v = pick(*s->fl[comp], field);
    if (!v || v->size() < n) return 1;
    std::memcpy(dst, v->data(), n * sizeof(double));
    return 0;
}

int ref_solid_surface_force(void* h) {
    RefSys* s = static_cast<RefSys*>(h);
    try {
        s->lbm->compute_solid_surface_force(*s->geom, *s->fl[0], *s->fl[1]);
    } catch (...) {
        return 1;
    }
    return 0;
}

int ref_run(void* h, int nsteps, const double* force0, const double* force1) {
    RefSys* s = static_cast<RefSys*>(h);
    std::vector<double> f0, f1;
    if (force0) f0.assign(force0, force0 + 9);
    if (force1) f1.assign(force1, force1 + 9);
    try {
        for (int t = 0; t < nsteps; ++t) {
            if (s->ncomp == 1) {
                s->lbm->collide(*s->geom, *s->fl[0]);
                if (force0) s->lbm->add_volume_force(*s->geom, *s->fl[0], f0);
                s->lbm->stream(*s->geom, *s->fl[0]);
            } else {
                s->fl[0]->compute_density();
                s->fl[1]->compute_density();
                s->lbm->compute_fluid_repulsive_interactions(*s->geom, *s->fl[0], *s->fl[1]);
                s->lbm->compute_equilibrium_velocities(*s->geom, *s->fl[0], *s->fl[1]);
                s->lbm->collide(*s->fl[0], *s->fl[1]);
                if (force0) s->lbm->add_volume_force(*s->geom, *s->fl[0], f0);
                if (force1) s->lbm->add_volume_force(*s->geom, *s->fl[1], f1);
                s->lbm->stream(*s->geom, *s->fl[0], *s->fl[1]);
            }
        }
    } catch (...) {
        return 1;
    }
    return 0;
}

/* Geometry(nx, ny) + add_walls(dh, "x"/"y") (skipped when dh == 0) + a list
 * of add_circle calls; the resulting mask is written to out_mask. */
int ref_geometry(int nx, int ny, size_t dh, int where_x, int ncirc, const size_t* d,
                 const size_t* xc, const size_t* yc, int* out_mask) {
    try {
        Geometry g((size_t)nx, (size_t)ny);
        if (dh > 0) g.add_walls(dh, where_x ? "x" : "y");
        for (int k = 0; k < ncirc; ++k) g.add_circle(d[k], xc[k], yc[k]);
        const std::vector<int>& m = g.get_geom();
        std::memcpy(out_mask, m.data(), m.size() * sizeof(int));
    } catch (const std::invalid_argument&) {
        return 2;
    } catch (...) {
        return 1;
    }
    return 0;
}

int ref_add_rectangle(int nx, int ny, size_t lx, size_t ly, size_t xc, size_t yc, int* out_mask) {
    try {
        Geometry g((size_t)nx, (size_t)ny);
        g.add_rectangle(lx, ly, xc, yc);
        const std::vector<int>& m = g.get_geom();
        std::memcpy(out_mask, m.data(), m.size() * sizeof(int));
    } catch (...) {
        return 1;
    }
    return 0;
}

/* Geometry(nx, ny) + one object call: kind 0 add_ellipse(p0, p1, p2, p3),
 * 1 add_square(p0, p1, p2), 2 add_array by gaps, 3 add_array by numbers
 * (object type `type`, nprops properties, bounds {x0, x1, y0, y1}, then the
 * two gaps or counts u, v).  Returns 0, or 1 runtime_error, 2 invalid_argument,
 * 3 out_of_range, 4 another exception; the mask is written only on success. */
int ref_geometry_object(int nx, int ny, int kind, const char* type, int nprops, const size_t* p,
                        const size_t* bounds, size_t u, size_t v, int* out_mask) {
    try {
        Geometry g((size_t)nx, (size_t)ny);
        std::vector<size_t> props(p, p + nprops);
        std::vector<std::vector<size_t>> b = {{bounds[0], bounds[1]}, {bounds[2], bounds[3]}};
        if (kind == 0) g.add_ellipse(p[0], p[1], p[2], p[3]);
        else if (kind == 1) g.add_square(p[0], p[1], p[2]);
        else if (kind == 2) g.add_array(props, b, u, v, type);
        else g.add_array(props, b, std::vector<size_t>{u, v}, type);
        const std::vector<int>& m = g.get_geom();
        std::memcpy(out_mask, m.data(), m.size() * sizeof(int));
    } catch (const std::invalid_argument&) {
        return 2;
    } catch (const std::out_of_range&) {
        return 3;
    } catch (const std::runtime_error&) {
        return 1;
    } catch (...) {
        return 4;
    }
    return 0;
}

/* Fluid::write_f(prefix) of component c: the reference's 6-digit text files
 * prefix_0.txt .. prefix_8.txt (src/fluid.cpp:346-353). */
int ref_write_f(void* h, int c, const char* prefix) {
    RefSys* s = static_cast<RefSys*>(h);
    try {
        s->fl[c]->write_f(prefix);
    } catch (...) {
        return 1;
    }
    return 0;
}

int ref_init_droplet(void* h, double rb, double rd, double rbd, double rdb,
                     double xc, double yc, double r) {
    RefSys* s = static_cast<RefSys*>(h);
    try {
        s->lbm->initialize_droplet(*s->geom, *s->fl[0], *s->fl[1], rb, rd, rbd, rdb, xc, yc, r);
    } catch (...) {
        return 1;
    }
    return 0;
}

int ref_init_rectangle(void* h, double rb, double rd, double rbd, double rdb,
                       double xc, double yc, double hx, double hy) {
    RefSys* s = static_cast<RefSys*>(h);
    try {
        s->lbm->initialize_fluid_rectangle(*s->geom, *s->fl[0], *s->fl[1], rb, rd, rbd, rdb,
                                           xc, yc, hx, hy);
    } catch (...) {
        return 1;
    }
    return 0;
}

}  // extern "C"
