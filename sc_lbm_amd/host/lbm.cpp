// lbm.cpp -- LBM drop-in (see include/lbm.h) and the operator recorder.
//
// Every device operator of the reference API is appended to a per-thread
// list instead of running at once.  stream() closes a time step: when the
// list holds exactly the reference's step loop (the sequences documented in
// include/lbm.h) it becomes a fused step (scl_step1 / scl_step2), which
// reads and writes the populations once; otherwise each recorded operator
// runs as its own kernel, in order.  Fused steps are not issued at once:
// consecutive identical ones (same fluids, geometry and forces) are counted
// and issued as one scl_step1/2(n) when anything else needs the device or a
// host array (any other operator, a host read, sclbm_flush, a geometry or
// fluid change), so the deviation forms reach their multi-step kernels.  Any
// host read runs what is pending first, so results never depend on the
// choice.
#include <algorithm>
#include <stdexcept>
#include <string>

#include "engine.h"
#include "lbm.h"

namespace sclbm_detail {

struct Op {
    int kind;
    Fluid* a;
    Fluid* b;
    const Geometry* g;
    std::vector<double> force;
};

// A run of identical fused steps not yet issued.
struct Pending {
    int n = 0;   // steps (0: none)
    Fluid* a = nullptr;
    Fluid* b = nullptr;   // null: one component
    scl_geometry* g = nullptr;
    bool has_fa = false, has_fb = false;
    std::vector<double> fa, fb;
    bool same(const Pending& o) const {
        return a == o.a && b == o.b && g == o.g && has_fa == o.has_fa && has_fb == o.has_fb && fa == o.fa &&
               fb == o.fb;
    }
};

struct Engine {
    static thread_local std::vector<Op> log;
    static thread_local bool running;
    static thread_local Pending pend;

    static void prepare(Fluid* f, const Geometry* g) {
        if (!f->dev) {
            const Geometry* geom = g ? g : f->dev_geom;
            if (!geom) throw std::logic_error("Fluid " + f->name + " is not initialised (no *_ini call)");
            f->device_field(*geom);
        }
        f->push_device();
    }

    static scl_geometry* geom_of(const Op& op) { return op.g ? op.g->device() : nullptr; }

    static void run_one(const Op& op);
    static bool try_fused();
    static void defer(Pending c);
    static void run_pending();
    static void flush();
};

thread_local std::vector<Op> Engine::log;
thread_local bool Engine::running = false;
thread_local Pending Engine::pend;

void raise_status(int rc) {
    if (rc == SCL_OK) return;
    const std::string msg = std::string("sclbm: ") + scl_last_error();
    if (rc == SCL_ERR_RANGE) throw std::out_of_range(msg);
    if (rc == SCL_ERR_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

void Engine::run_one(const Op& op) {
    Fluid* a = op.a;
    Fluid* b = op.b;
    prepare(a, op.g);
    if (b) prepare(b, op.g);
    scl_geometry* g = geom_of(op);
    switch (op.kind) {
        case OP_DENSITY: raise_status(scl_compute_density(a->dev)); break;
        case OP_VELOCITIES: raise_status(scl_compute_velocities(g, a->dev)); break;
        case OP_FEQ_GEOM: raise_status(scl_compute_f_equilibrium_geom(g, a->dev)); break;
        case OP_FEQ_MC: raise_status(scl_compute_f_equilibrium(a->dev)); break;
        case OP_REPULSION: raise_status(scl_fluid_repulsion(g, a->dev, b->dev)); break;
        case OP_EQVEL: raise_status(scl_equilibrium_velocities(g, a->dev, b->dev)); break;
        case OP_COLLIDE1: raise_status(scl_collide1(g, a->dev)); break;
        case OP_COLLIDE2: raise_status(scl_collide2(a->dev, b->dev)); break;
        case OP_VF:
            raise_status(scl_volume_force(g, a->dev, op.force.data()));
            if (b) raise_status(scl_volume_force(g, b->dev, op.force.data()));
            break;
        case OP_STREAM1: raise_status(scl_stream1(g, a->dev)); break;
        case OP_STREAM2: raise_status(scl_stream2(g, a->dev, b->dev)); break;
        default: throw std::logic_error("unknown recorded operator");
    }
    a->device_wrote();
    if (b) b->device_wrote();
}

// The recorded list is one reference time step that the fused kernel can run.
bool Engine::try_fused() {
    if (log.empty()) return false;
    const Op& last = log.back();
    if (last.kind == OP_STREAM1) {
        Fluid* f = last.a;
        if (f->precision != SCL_FP64 && log.size() > 3) return false;
        size_t k = 0;
        if (log.size() < 2 || log.size() > 3) return false;
        if (log[k].kind != OP_COLLIDE1 || log[k].a != f) return false;
        ++k;
        const std::vector<double>* force = nullptr;
        if (log[k].kind == OP_VF) {
            if (log[k].a != f || log[k].b) return false;
            force = &log[k].force;
            ++k;
        }
        if (k != log.size() - 1) return false;
        const Geometry* g = last.g;
        if (log[0].g != g) return false;
        Pending c;
        c.a = f;
        c.g = g->device();
        if (force) {
            c.has_fa = true;
            c.fa = *force;
        }
        if (!(pend.n > 0 && pend.same(c))) {
            run_pending();
            prepare(f, g);
        }
        defer(std::move(c));
        return true;
    }
    if (last.kind != OP_STREAM2) return false;
    Fluid* A = last.a;
    Fluid* B = last.b;
    const Geometry* g = last.g;
    if (!A->solid_force_from_mask || !B->solid_force_from_mask) return false;
    size_t k = 0;
    bool da = false, db = false;
    while (k < log.size() && log[k].kind == OP_DENSITY) {
        if (log[k].a == A && !da) da = true;
        else if (log[k].a == B && !db) db = true;
        else return false;
        ++k;
    }
    if (!da || !db) return false;
    auto is = [&](size_t i, int kind) {
        return i < log.size() && log[i].kind == kind && log[i].a == A && log[i].b == B;
    };
    if (!is(k, OP_REPULSION) || log[k].g != g) return false;
    ++k;
    if (!is(k, OP_EQVEL) || log[k].g != g) return false;
    ++k;
    if (!is(k, OP_COLLIDE2)) return false;
    ++k;
    const std::vector<double>* fa = nullptr;
    const std::vector<double>* fb = nullptr;
    while (k < log.size() && log[k].kind == OP_VF) {
        const Op& v = log[k];
        if (v.g != g) return false;
        if (v.b) {   // the two-fluid overload applies one vector to both
            if (!((v.a == A && v.b == B) || (v.a == B && v.b == A)) || fa || fb) return false;
            fa = fb = &v.force;
        } else if (v.a == A && !fa) {
            fa = &v.force;
        } else if (v.a == B && !fb) {
            fb = &v.force;
        } else {
            return false;
        }
        ++k;
    }
    if (k != log.size() - 1) return false;
    Pending c;
    c.a = A;
    c.b = B;
    c.g = g->device();
    if (fa) {
        c.has_fa = true;
        c.fa = *fa;
    }
    if (fb) {
        c.has_fb = true;
        c.fb = *fb;
    }
    if (!(pend.n > 0 && pend.same(c))) {
        run_pending();
        prepare(A, g);
        prepare(B, g);
    }
    defer(std::move(c));
    return true;
}

// Counts a fused step onto the pending run (the same step as the run's) or
// issues the run and starts a new one.  The fluids' host mirrors are stale
// from here on, so a host read runs the steps first.
void Engine::defer(Pending c) {
    if (pend.n > 0 && pend.same(c)) {
        ++pend.n;
    } else {
        run_pending();
        pend = std::move(c);
        pend.n = 1;
    }
    pend.a->device_wrote();
    if (pend.b) pend.b->device_wrote();
}

void Engine::run_pending() {
    if (pend.n == 0) return;
    Pending p = std::move(pend);
    pend = Pending();
    const double* fa = p.has_fa ? p.fa.data() : nullptr;
    const double* fb = p.has_fb ? p.fb.data() : nullptr;
    const int rc = p.b ? scl_step2(p.g, p.a->dev, p.b->dev, fa, fb, p.n) : scl_step1(p.g, p.a->dev, fa, p.n);
    raise_status(rc);
    p.a->device_wrote();
    if (p.b) p.b->device_wrote();
}

void Engine::flush() {
    if (running || (log.empty() && pend.n == 0)) return;
    running = true;
    try {
        run_pending();
        std::vector<Op> ops;
        ops.swap(log);
        for (const Op& op : ops) run_one(op);
    } catch (...) {
        running = false;
        throw;
    }
    running = false;
}

void flush_log() { Engine::flush(); }

void forget(const Fluid* f) {
    if (Engine::running) return;
    if (Engine::pend.n > 0 && (Engine::pend.a == f || Engine::pend.b == f)) {
        try {
            Engine::flush();
        } catch (...) {
            Engine::pend = Pending();
            Engine::log.clear();
        }
        return;
    }
    for (const Op& op : Engine::log)
        if (op.a == f || op.b == f) {
            try {
                Engine::flush();
            } catch (...) {
                Engine::log.clear();
            }
            return;
        }
}

void record(Op op) { Engine::log.push_back(std::move(op)); }

void record_fluid_op(int kind, Fluid* f, const Geometry* g) { record(Op{kind, f, nullptr, g, {}}); }

void close_step() {
    if (Engine::running) return;
    bool fused = false;
    try {
        fused = Engine::try_fused();
    } catch (...) {
        Engine::log.clear();
        throw;
    }
    if (fused) Engine::log.clear();
    else Engine::flush();
}

}  // namespace sclbm_detail

using sclbm_detail::Op;
using sclbm_detail::record;

void sclbm_flush() { sclbm_detail::flush_log(); }

LBM::LBM(const Geometry& geom) {
    Nx = geom.Nx();
    Ny = geom.Ny();
    Ntot = Nx * Ny;
}

// src/lbm.cpp:11-48: f_j = rho/9 in all directions on fluid sites; inside
// the circle (double coordinates) the droplet densities.
void LBM::initialize_droplet(const Geometry& geom, Fluid& bulk, Fluid& droplet, const double rho_bulk,
                             const double rho_droplet, const double rho_b_in_d, const double rho_d_in_b,
                             const double xc, const double yc, const double radius) {
    std::vector<double>& fb = bulk.get_f_dist();
    std::vector<double>& fd = droplet.get_f_dist();
    for (size_t ai = 0; ai < Ntot; ++ai) {
        const int xi = (int)(ai % Nx);
        const int yj = (int)(((ai - xi) / Nx) % Ny);
        if (geom(ai) == 0) continue;
        const double dx = static_cast<double>(xi) - xc, dy = static_cast<double>(yj) - yc;
        const bool inside = (dx * dx + dy * dy) <= radius * radius;
        for (size_t dj = 0; dj < Ndir; ++dj) {
            fb.at(ai + dj * Ntot) = (inside ? rho_b_in_d : rho_bulk) / Ndir;
            fd.at(ai + dj * Ntot) = (inside ? rho_droplet : rho_d_in_b) / Ndir;
        }
    }
}

// src/lbm.cpp:51-99: bulk everywhere on fluid sites, then the rectangle
// x in [int(xc-hx+1), int(xc+hx-1)], y in [int(yc-hy+1), int(yc+hy-1)].
void LBM::initialize_fluid_rectangle(const Geometry& geom, Fluid& bulk, Fluid& droplet, const double rho_bulk,
                                     const double rho_droplet, const double rho_b_in_d, const double rho_d_in_b,
                                     const double xc, const double yc, const double half_Lx,
                                     const double half_Ly) {
    std::vector<double>& fb = bulk.get_f_dist();
    std::vector<double>& fd = droplet.get_f_dist();
    for (size_t ai = 0; ai < Ntot; ++ai) {
        if (geom(ai) == 0) continue;
        for (size_t dj = 0; dj < Ndir; ++dj) {
            fb.at(ai + dj * Ntot) = rho_bulk / Ndir;
            fd.at(ai + dj * Ntot) = rho_d_in_b / Ndir;
        }
    }
    const int x0 = xc - half_Lx + 1, xf = xc + half_Lx - 1;
    const int y0 = yc - half_Ly + 1, yf = yc + half_Ly - 1;
    for (int xi = x0; xi <= xf; ++xi)
        for (int yj = y0; yj <= yf; ++yj) {
            if (geom(xi, yj) == 0) continue;
            for (size_t dj = 0; dj < Ndir; ++dj) {
                const size_t ijk = static_cast<size_t>(dj * Ntot + yj * Nx + xi);
                fb.at(ijk) = rho_b_in_d / Ndir;
                fd.at(ijk) = rho_droplet / Ndir;
            }
        }
}

// src/lbm.cpp:102-141 with src/fluid.cpp:154-160.  Runs at once (it is
// called once per run).  The host mirrors grow by Ntot entries, as the
// reference appends; the device keeps the first set.
void LBM::compute_solid_surface_force(const Geometry& geom, Fluid& f1, Fluid& f2) {
    sclbm_flush();
    for (Fluid* f : {&f1, &f2}) {
        sclbm_detail::Engine::prepare(f, &geom);
        f->sync_host(Fluid::A_FSX);
        f->sync_host(Fluid::A_FSY);
    }
    const bool first1 = f1.arr[Fluid::A_FSX].empty(), first2 = f2.arr[Fluid::A_FSX].empty();
    // the values a repeated call appends (kept on the host only)
    std::vector<double> sx(Ntot, 0.0), sy(Ntot, 0.0);
    if (!first1 || !first2) {
        static const int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
        static const int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
        static const double w[9] = {0.0, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
        for (size_t ai = 0; ai < Ntot; ++ai) {
            const int xi = (int)(ai % Nx), yj = (int)(ai / Nx);
            if (geom(ai) == 0) continue;
            for (int d = 1; d < 9; ++d) {
                const int xn = (xi + cx[d] + (int)Nx) % (int)Nx, yn = (yj + cy[d] + (int)Ny) % (int)Ny;
                if (geom(xn, yn) == 0) {
                    sx[ai] += w[d] * cx[d];
                    sy[ai] += w[d] * cy[d];
                }
            }
        }
    }
    sclbm_detail::raise_status(scl_solid_surface_force(geom.device(), f1.dev, f2.dev));
    for (Fluid* f : {&f1, &f2}) {
        const bool first = f->arr[Fluid::A_FSX].empty();
        if (first) {
            f->arr[Fluid::A_FSX].assign(Ntot, 0.0);
            f->arr[Fluid::A_FSY].assign(Ntot, 0.0);
            f->host_ok[Fluid::A_FSX] = false;
            f->host_ok[Fluid::A_FSY] = false;
            f->dev_ok[Fluid::A_FSX] = true;
            f->dev_ok[Fluid::A_FSY] = true;
            f->solid_force_from_mask = true;
        } else {
            for (size_t i = 0; i < Ntot; ++i) {
                f->arr[Fluid::A_FSX].push_back(-1.0 * f->Gsolid * sx[i]);
                f->arr[Fluid::A_FSY].push_back(-1.0 * f->Gsolid * sy[i]);
            }
        }
        f->solid_force_set = true;
    }
}

void LBM::compute_fluid_repulsive_interactions(const Geometry& geom, Fluid& f1, Fluid& f2) {
    if (f1.arr[Fluid::A_FRX].size() < Ntot || f2.arr[Fluid::A_FRX].size() < Ntot)
        throw std::out_of_range("vector::_M_range_check: repulsive force arrays are not initialised");
    record(Op{sclbm_detail::OP_REPULSION, &f1, &f2, &geom, {}});
}

void LBM::compute_equilibrium_velocities(Geometry& geom, Fluid& f1, Fluid& f2) {
    for (Fluid* f : {&f1, &f2})
        if (f->arr[Fluid::A_UEQX].size() < Ntot || f->arr[Fluid::A_FRX].size() < Ntot ||
            f->arr[Fluid::A_FSX].size() < Ntot)
            throw std::out_of_range("vector::_M_range_check: equilibrium velocity inputs are not initialised");
    record(Op{sclbm_detail::OP_EQVEL, &f1, &f2, &geom, {}});
}

void LBM::collide(const Geometry& geom, Fluid& f1) { record(Op{sclbm_detail::OP_COLLIDE1, &f1, nullptr, &geom, {}}); }

void LBM::collide(Fluid& f1, Fluid& f2) { record(Op{sclbm_detail::OP_COLLIDE2, &f1, &f2, nullptr, {}}); }

void LBM::add_volume_force(const Geometry& geom, Fluid& f1, const std::vector<double>& force) {
    if (force.size() < Ndir) throw std::out_of_range("vector::_M_range_check: force needs 9 entries");
    record(Op{sclbm_detail::OP_VF, &f1, nullptr, &geom, std::vector<double>(force.begin(), force.begin() + 9)});
}

void LBM::add_volume_force(const Geometry& geom, Fluid& f1, Fluid& f2, const std::vector<double>& force) {
    if (force.size() < Ndir) throw std::out_of_range("vector::_M_range_check: force needs 9 entries");
    record(Op{sclbm_detail::OP_VF, &f1, &f2, &geom, std::vector<double>(force.begin(), force.begin() + 9)});
}

void LBM::stream(const Geometry& geom, Fluid& f1) {
    record(Op{sclbm_detail::OP_STREAM1, &f1, nullptr, &geom, {}});
    sclbm_detail::close_step();
}

void LBM::stream(const Geometry& geom, Fluid& f1, Fluid& f2) {
    record(Op{sclbm_detail::OP_STREAM2, &f1, &f2, &geom, {}});
    sclbm_detail::close_step();
}
