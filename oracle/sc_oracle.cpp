/*
 * sc_oracle.cpp -- CPU oracle for the D2Q9 Shan-Chen step (test infrastructure).
 *
 * ORACLE: used only by tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py, as the checker.  Never linked into or called by the product.
 *
 * A serial restatement of the reference arithmetic, written to reproduce its
 * IEEE-754 double results bit for bit: the same operations in the same order,
 * no contraction into FMA (built with -ffp-contract=off), no reassociation.
 * Each routine cites the reference lines it follows.  The parity pins for each
 * routine are listed in tests/test_oracle.py (golden vectors written by the
 * compiled reference, the reference's closed-form fixtures, invariants).
 */
#include "sc_oracle.h"

#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

const int Q = 9;
/* Lattice velocities and opposites, include/lbm.h:121-128. */
const int CX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
const int CY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
const int OPP[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
/* Interaction weights (solid and fluid-fluid use the same set),
 * include/lbm.h:118-122. */
const double WINT[Q] = {0.0, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                        1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
/* Equilibrium constants, include/fluid.h:282-283. */
const double W0 = 4.0 / 9.0, W1 = 1.0 / 9.0, W2 = 1.0 / 36.0;
const double K1 = 3.0, K2 = 4.5, K3 = 1.5;

struct Comp {
    double tau = 1.0, omega = 1.0;
    double g_fluid = 0.0, g_solid = 0.0;
    bool has_solid_force = false;
    std::vector<double> f, feq, rho, ux, uy, ueqx, ueqy, frx, fry, fsx, fsy;
};

}  // namespace

struct oracle_sys {
    int nx = 0, ny = 0, ncomp = 0;
    size_t n = 0;
    std::vector<int> mask;
    Comp c[2];
    std::vector<double> tmp[2];

    bool fluid(int x, int y) const { return mask[(size_t)y * nx + x] == 1; }
    int wx(int x) const { return x < 0 ? nx - 1 : (x >= nx ? 0 : x); }
    int wy(int y) const { return y < 0 ? ny - 1 : (y >= ny ? 0 : y); }
};

static std::vector<double>* field_ptr(oracle_sys* s, int comp, int field) {
    if (!s || comp < 0 || comp >= s->ncomp) return nullptr;
    Comp& c = s->c[comp];
    switch (field) {
        case ORACLE_F: return &c.f;
        case ORACLE_FEQ: return &c.feq;
        case ORACLE_RHO: return &c.rho;
        case ORACLE_UX: return &c.ux;
        case ORACLE_UY: return &c.uy;
        case ORACLE_UEQX: return &c.ueqx;
        case ORACLE_UEQY: return &c.ueqy;
        case ORACLE_FRX: return &c.frx;
        case ORACLE_FRY: return &c.fry;
        case ORACLE_FSX: return &c.fsx;
        case ORACLE_FSY: return &c.fsy;
        default: return nullptr;
    }
}

extern "C" {

oracle_sys* oracle_create(int nx, int ny, int ncomp, const int* mask,
                          const double* tau, const double* g_fluid,
                          const double* g_solid) {
    if (nx <= 0 || ny <= 0 || (ncomp != 1 && ncomp != 2) || !mask || !tau) return nullptr;
    oracle_sys* s = new oracle_sys;
    s->nx = nx;
    s->ny = ny;
    s->ncomp = ncomp;
    s->n = (size_t)nx * ny;
    s->mask.assign(mask, mask + s->n);
    for (int k = 0; k < ncomp; ++k) {
        Comp& c = s->c[k];
        c.tau = tau[k];
        c.omega = 1 / c.tau;                       /* include/fluid.h:62 */
        c.g_fluid = g_fluid ? g_fluid[k] : 0.0;
        c.g_solid = g_solid ? g_solid[k] : 0.0;
        c.f.assign(Q * s->n, 0.0);
        c.feq.assign(Q * s->n, 0.0);
        for (auto* v : {&c.rho, &c.ux, &c.uy, &c.ueqx, &c.ueqy, &c.frx, &c.fry, &c.fsx, &c.fsy})
            v->assign(s->n, 0.0);
        s->tmp[k].assign(Q * s->n, 0.0);
    }
    return s;
}

void oracle_destroy(oracle_sys* s) { delete s; }

int oracle_set(oracle_sys* s, int comp, int field, const double* src, size_t n) {
    std::vector<double>* v = field_ptr(s, comp, field);
    if (!v || v->size() != n || !src) return 1;
    std::copy(src, src + n, v->begin());
    return 0;
}

int oracle_get(const oracle_sys* s, int comp, int field, double* dst, size_t n) {
    std::vector<double>* v = field_ptr(const_cast<oracle_sys*>(s), comp, field);
    if (!v || v->size() != n || !dst) return 1;
    std::copy(v->begin(), v->end(), dst);
    return 0;
}

/* Fluid-solid force, src/lbm.cpp:102-141: for each fluid site, the weighted
 * lattice vectors pointing at solid neighbours (periodic), summed in
 * direction order 1..8 from 0.0; then each fluid scales it by (-1.0*Gs),
 * src/fluid.cpp:154-160.  Only the first call takes effect (the reference
 * appends to F_solid and reads the first N entries). */
void oracle_solid_surface_force(oracle_sys* s) {
    const int nx = s->nx, ny = s->ny;
    std::vector<double> sx(s->n, 0.0), sy(s->n, 0.0);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            if (!s->fluid(x, y)) continue;
            size_t at = (size_t)y * nx + x;
            for (int i = 1; i < Q; ++i) {
                if (s->fluid(s->wx(x + CX[i]), s->wy(y + CY[i]))) continue;
                sx[at] += WINT[i] * CX[i];
                sy[at] += WINT[i] * CY[i];
            }
        }
    for (int k = 0; k < s->ncomp; ++k) {
        Comp& c = s->c[k];
        if (c.has_solid_force) continue;
        const double gneg = -1.0 * c.g_solid;
        for (size_t a = 0; a < s->n; ++a) {
            c.fsx[a] = gneg * sx[a];
            c.fsy[a] = gneg * sy[a];
        }
        c.has_solid_force = true;
    }
}

/* src/fluid.cpp:175-183: rho = sum_j f_j from 0.0 in j order, every site. */
void oracle_density(oracle_sys* s, int comp) {
    Comp& c = s->c[comp];
    for (size_t a = 0; a < s->n; ++a) {
        double r = 0.0;
        for (int j = 0; j < Q; ++j) r += c.f[j * s->n + a];
        c.rho[a] = r;
    }
}

/* src/fluid.cpp:186-200: u = (sum_j f_j c_j)/rho on fluid sites, 0 elsewhere. */
void oracle_velocities(oracle_sys* s, int comp) {
    Comp& c = s->c[comp];
    for (size_t a = 0; a < s->n; ++a) {
        double jx = 0.0, jy = 0.0;
        if (s->mask[a] == 1) {
            for (int j = 0; j < Q; ++j) {
                jx += c.f[j * s->n + a] * (double)CX[j];
                jy += c.f[j * s->n + a] * (double)CY[j];
            }
            jx /= c.rho[a];
            jy /= c.rho[a];
        }
        c.ux[a] = jx;
        c.uy[a] = jy;
    }
}

/* Second-order equilibrium of one site, src/fluid.cpp:217-242 / 254-276,
 * evaluated left to right exactly as written there. */
static inline void feq_site(double rho, double ux, double uy, double* e) {
    const double r0 = W0 * rho, r1 = W1 * rho, r2 = W2 * rho;
    const double uxx = ux * ux, uyy = uy * uy;
    const double a5 = ux + uy, a6 = -ux + uy, a7 = -ux - uy, a8 = ux - uy;
    const double usq = uxx + uyy;
    e[0] = r0 * (1.0 - K3 * usq);
    e[1] = r1 * (1.0 + K1 * ux + K2 * uxx - K3 * usq);
    e[2] = r1 * (1.0 + K1 * uy + K2 * uyy - K3 * usq);
    e[3] = r1 * (1.0 - K1 * ux + K2 * uxx - K3 * usq);
    e[4] = r1 * (1.0 - K1 * uy + K2 * uyy - K3 * usq);
    e[5] = r2 * (1.0 + K1 * a5 + K2 * a5 * a5 - K3 * usq);
    e[6] = r2 * (1.0 + K1 * a6 + K2 * a6 * a6 - K3 * usq);
    e[7] = r2 * (1.0 + K1 * a7 + K2 * a7 * a7 - K3 * usq);
    e[8] = r2 * (1.0 + K1 * a8 + K2 * a8 * a8 - K3 * usq);
}

static void feq_from(oracle_sys* s, Comp& c, const std::vector<double>& ux,
                     const std::vector<double>& uy) {
    double e[Q];
    for (size_t a = 0; a < s->n; ++a) {
        feq_site(c.rho[a], ux[a], uy[a], e);
        for (int j = 0; j < Q; ++j) c.feq[j * s->n + a] = e[j];
    }
}

/* src/fluid.cpp:210-243: density, velocity, then feq on every site. */
void oracle_feq_geom(oracle_sys* s, int comp) {
    oracle_density(s, comp);
    oracle_velocities(s, comp);
    feq_from(s, s->c[comp], s->c[comp].ux, s->c[comp].uy);
}

/* src/fluid.cpp:246-278: feq from rho and u_eq, every site. */
void oracle_feq_mc(oracle_sys* s, int comp) {
    Comp& c = s->c[comp];
    feq_from(s, c, c.ueqx, c.ueqy);
}

/* src/lbm.cpp:144-206: psi = rho; for each fluid site the 8-neighbour sum
 * over fluid neighbours of (w_i*c_i)*psi_other, in direction order from 0.0,
 * then scaled by (-1.0*G_self)*psi_self.  Solid sites keep 0. */
void oracle_repulsion(oracle_sys* s) {
    Comp& A = s->c[0];
    Comp& B = s->c[1];
    std::fill(A.frx.begin(), A.frx.end(), 0.0);
    std::fill(A.fry.begin(), A.fry.end(), 0.0);
    std::fill(B.frx.begin(), B.frx.end(), 0.0);
    std::fill(B.fry.begin(), B.fry.end(), 0.0);
    const double ga = -1.0 * A.g_fluid, gb = -1.0 * B.g_fluid;
    const int nx = s->nx, ny = s->ny;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            if (!s->fluid(x, y)) continue;
            size_t at = (size_t)y * nx + x;
            double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
            for (int i = 1; i < Q; ++i) {
                int xn = s->wx(x + CX[i]), yn = s->wy(y + CY[i]);
                if (!s->fluid(xn, yn)) continue;
                size_t nb = (size_t)yn * nx + xn;
                ax += WINT[i] * CX[i] * B.rho[nb];
                ay += WINT[i] * CY[i] * B.rho[nb];
                bx += WINT[i] * CX[i] * A.rho[nb];
                by += WINT[i] * CY[i] * A.rho[nb];
            }
            ax *= ga * A.rho[at];
            ay *= ga * A.rho[at];
            bx *= gb * B.rho[at];
            by *= gb * B.rho[at];
            A.frx[at] = ax; A.fry[at] = ay;
            B.frx[at] = bx; B.fry[at] = by;
        }
}

/* The reference's near-zero test, include/utils.h:145-150, with num2 = 0. */
static inline bool is_zero_density(double r) {
    const double m = std::max({1.0, std::fabs(r), 0.0});
    return std::fabs(r - 0.0) <= 1e-16 * m;
}

/* src/lbm.cpp:209-272: momentum (stored in ux/uy), composite velocity
 * weighted by omega, equilibrium velocity per component with the force
 * shift; a component whose density is zero keeps its previous u_eq. */
void oracle_eq_velocities(oracle_sys* s) {
    Comp& A = s->c[0];
    Comp& B = s->c[1];
    if (!A.has_solid_force || !B.has_solid_force) return;  /* reference throws (.at on empty) */
    const double oa = A.omega, ob = B.omega;
    const double ia = 1.0 / oa, ib = 1.0 / ob;
    const size_t n = s->n;
    std::fill(A.ux.begin(), A.ux.end(), 0.0);
    std::fill(A.uy.begin(), A.uy.end(), 0.0);
    std::fill(B.ux.begin(), B.ux.end(), 0.0);
    std::fill(B.uy.begin(), B.uy.end(), 0.0);
    for (size_t a = 0; a < n; ++a) {
        if (s->mask[a] != 1) continue;
        double jax = 0.0, jay = 0.0, jbx = 0.0, jby = 0.0;
        for (int j = 0; j < Q; ++j) {
            jax += A.f[j * n + a] * CX[j];
            jay += A.f[j * n + a] * CY[j];
            jbx += B.f[j * n + a] * CX[j];
            jby += B.f[j * n + a] * CY[j];
        }
        A.ux[a] = jax; A.uy[a] = jay;
        B.ux[a] = jbx; B.uy[a] = jby;
        const double ucx = (jax * oa + jbx * ob) / (A.rho[a] * oa + B.rho[a] * ob);
        const double ucy = (jay * oa + jby * ob) / (A.rho[a] * oa + B.rho[a] * ob);
        if (!is_zero_density(A.rho[a])) {
            A.ueqx[a] = ucx + A.frx[a] * ia / A.rho[a] + A.fsx[a] * ia;
            A.ueqy[a] = ucy + A.fry[a] * ia / A.rho[a] + A.fsy[a] * ia;
        }
        if (!is_zero_density(B.rho[a])) {
            B.ueqx[a] = ucx + B.frx[a] * ib / B.rho[a] + B.fsx[a] * ib;
            B.ueqy[a] = ucy + B.fry[a] * ib / B.rho[a] + B.fsy[a] * ib;
        }
    }
}

static void bgk(oracle_sys* s, Comp& c) {
    const double w = c.omega;
    for (size_t k = 0; k < Q * s->n; ++k) c.f[k] = (1.0 - w) * c.f[k] + w * c.feq[k];
}

/* src/lbm.cpp:275-288 */
void oracle_collide1(oracle_sys* s) {
    oracle_feq_geom(s, 0);
    bgk(s, s->c[0]);
}

/* src/lbm.cpp:291-312 */
void oracle_collide2(oracle_sys* s) {
    oracle_feq_mc(s, 0);
    oracle_feq_mc(s, 1);
    bgk(s, s->c[0]);
    bgk(s, s->c[1]);
}

/* src/lbm.cpp:315-340: f_j += force_j on fluid sites. */
void oracle_volume_force(oracle_sys* s, int comp, const double* force) {
    Comp& c = s->c[comp];
    for (size_t a = 0; a < s->n; ++a) {
        if (s->mask[a] != 1) continue;
        for (int j = 0; j < Q; ++j) c.f[j * s->n + a] += force[j];
    }
}

/* src/lbm.cpp:343-435: push streaming from fluid sites with periodic wrap;
 * a population whose destination is solid is reflected into the opposite
 * direction of its own site (link bounce-back).  The result replaces f and
 * every slot not written (all solid sites) is zero. */
void oracle_stream(oracle_sys* s) {
    const int nx = s->nx, ny = s->ny;
    const size_t n = s->n;
    for (int k = 0; k < s->ncomp; ++k) {
        std::vector<double>& src = s->c[k].f;
        std::vector<double>& dst = s->tmp[k];
        std::fill(dst.begin(), dst.end(), 0.0);
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                if (!s->fluid(x, y)) continue;
                size_t at = (size_t)y * nx + x;
                dst[at] = src[at];
                for (int i = 1; i < Q; ++i) {
                    int xd = s->wx(x + CX[i]), yd = s->wy(y + CY[i]);
                    if (s->fluid(xd, yd))
                        dst[i * n + (size_t)yd * nx + xd] = src[i * n + at];
                    else
                        dst[OPP[i] * n + at] = src[i * n + at];
                }
            }
        src.swap(dst);
    }
}

void oracle_run(oracle_sys* s, int nsteps, const double* force0, const double* force1) {
    for (int t = 0; t < nsteps; ++t) {
        if (s->ncomp == 1) {
            oracle_collide1(s);
            if (force0) oracle_volume_force(s, 0, force0);
            oracle_stream(s);
        } else {
            oracle_density(s, 0);
            oracle_density(s, 1);
            oracle_repulsion(s);
            oracle_eq_velocities(s);
            oracle_collide2(s);
            if (force0) oracle_volume_force(s, 0, force0);
            if (force1) oracle_volume_force(s, 1, force1);
            oracle_stream(s);
        }
    }
}

void oracle_moments(const oracle_sys* s, int comp, double out[3]) {
    const Comp& c = s->c[comp];
    long double m = 0, jx = 0, jy = 0;
    for (int j = 0; j < Q; ++j)
        for (size_t a = 0; a < s->n; ++a) {
            long double v = c.f[j * s->n + a];
            m += v;
            jx += v * CX[j];
            jy += v * CY[j];
        }
    out[0] = (double)m;
    out[1] = (double)jx;
    out[2] = (double)jy;
}

/* src/geometry.cpp:34-69.  where_x != 0 means "x" (solid rows at the bottom
 * and top), 0 means "y" (solid columns at the left and right). */
int oracle_add_walls(int* mask, int nx, int ny, size_t dh, int where_x) {
    if (where_x) {
        if (dh > (size_t)ny) return 1;
        for (size_t r = 0; r < dh; ++r)
            for (int x = 0; x < nx; ++x) {
                mask[r * nx + x] = 0;
                mask[(ny - 1 - r) * (size_t)nx + x] = 0;
            }
    } else {
        if (dh > (size_t)nx) return 1;
        for (int y = 0; y < ny; ++y)
            for (size_t cidx = 0; cidx < dh; ++cidx) {
                mask[(size_t)y * nx + cidx] = 0;
                mask[(size_t)y * nx + (nx - 1 - cidx)] = 0;
            }
    }
    return 0;
}

/* Nominal bounds test of src/geometry.cpp:150-162 (uses the sizes before the
 * even-size decrement). */
static int object_bounds_ok(int nx, int ny, size_t lx, size_t ly, size_t xc, size_t yc) {
    if (((int)xc - (int)lx / 2) < 0 || ((int)yc - (int)ly / 2) < 0) return 0;
    if ((xc + lx / 2) >= (size_t)nx || (yc + ly / 2) >= (size_t)ny) return 0;
    return 1;
}

/* Node sets of src/geom_object/ellipse.cpp:12-33: even sizes shrink by one,
 * the box xc +- L/2, yc +- L/2 is scanned and a node is kept when the
 * integer test holds, evaluated in unsigned 64-bit arithmetic exactly as the
 * reference writes it (each product wraps mod 2^64, each division
 * truncates). */
int oracle_add_circle(int* mask, int nx, int ny, size_t d, size_t xc, size_t yc) {
    if (!object_bounds_ok(nx, ny, d, d, xc, yc)) return 1;
    size_t lx = d, ly = d;
    if (lx % 2 == 0) --lx;
    if (ly % 2 == 0) --ly;
    const size_t x0 = xc - lx / 2, x1 = xc + lx / 2;
    const size_t y0 = yc - ly / 2, y1 = yc + ly / 2;
    const size_t rhs = lx * lx * ly * ly / 16;
    std::vector<size_t> nodes;
    for (size_t ix = x0; ix <= x1; ++ix)
        for (size_t iy = y0; iy <= y1; ++iy) {
            const size_t dx = ix - xc, dy = iy - yc;
            const size_t lhs = dx * dx * ly * ly / 4 + dy * dy * lx * lx / 4;
            if (lhs <= rhs) nodes.push_back(iy * (size_t)nx + ix);
        }
    for (size_t v : nodes)
        if (v >= (size_t)nx * ny) return 1;  /* src/geometry.cpp:166-175 */
    for (size_t v : nodes) mask[v] = 0;
    return 0;
}

/* src/geom_object/rectangle.cpp:12-31 with the same bounds checks. */
int oracle_add_rectangle(int* mask, int nx, int ny, size_t lx, size_t ly, size_t xc, size_t yc) {
    if (!object_bounds_ok(nx, ny, lx, ly, xc, yc)) return 1;
    if (lx % 2 == 0) --lx;
    if (ly % 2 == 0) --ly;
    std::vector<size_t> nodes;
    for (size_t ix = xc - lx / 2; ix <= xc + lx / 2; ++ix)
        for (size_t iy = yc - ly / 2; iy <= yc + ly / 2; ++iy)
            nodes.push_back(iy * (size_t)nx + ix);
    for (size_t v : nodes)
        if (v >= (size_t)nx * ny) return 1;
    for (size_t v : nodes) mask[v] = 0;
    return 0;
}

/* src/lbm.cpp:11-48: f_j = rho/9 in every direction on fluid sites; the
 * droplet test is on double coordinates, (x-xc)^2 + (y-yc)^2 <= r^2. */
void oracle_init_droplet(oracle_sys* s, double rho_bulk, double rho_droplet,
                         double rho_b_in_d, double rho_d_in_b,
                         double xc, double yc, double radius) {
    const size_t n = s->n;
    Comp& bulk = s->c[0];
    Comp& drop = s->c[1];
    for (size_t a = 0; a < n; ++a) {
        if (s->mask[a] == 0) continue;
        const int x = (int)(a % s->nx), y = (int)(a / s->nx);
        const double dx = (double)x - xc, dy = (double)y - yc;
        const bool inside = (dx * dx + dy * dy) <= radius * radius;
        for (int j = 0; j < Q; ++j) {
            bulk.f[j * n + a] = (inside ? rho_b_in_d : rho_bulk) / 9;
            drop.f[j * n + a] = (inside ? rho_droplet : rho_d_in_b) / 9;
        }
    }
}

/* src/lbm.cpp:51-99: bulk everywhere on fluid sites, then the rectangle
 * x in [int(xc-hx+1), int(xc+hx-1)], y in [int(yc-hy+1), int(yc+hy-1)]
 * (double -> int conversion truncates). */
int oracle_init_rectangle(oracle_sys* s, double rho_bulk, double rho_droplet,
                          double rho_b_in_d, double rho_d_in_b,
                          double xc, double yc, double half_lx, double half_ly) {
    const size_t n = s->n;
    Comp& bulk = s->c[0];
    Comp& drop = s->c[1];
    for (size_t a = 0; a < n; ++a) {
        if (s->mask[a] == 0) continue;
        for (int j = 0; j < Q; ++j) {
            bulk.f[j * n + a] = rho_bulk / 9;
            drop.f[j * n + a] = rho_d_in_b / 9;
        }
    }
    const int x0 = (int)(xc - half_lx + 1), x1 = (int)(xc + half_lx - 1);
    const int y0 = (int)(yc - half_ly + 1), y1 = (int)(yc + half_ly - 1);
    for (int x = x0; x <= x1; ++x)
        for (int y = y0; y <= y1; ++y) {
            /* The reference converts the int coordinates to size_t and
             * bounds-checks only the flat index (geometry.h:241-243). */
            const size_t a = (size_t)(long long)y * (size_t)s->nx + (size_t)(long long)x;
            if (a >= n) return 1;
            if (s->mask[a] == 0) continue;
            for (int j = 0; j < Q; ++j) {
                bulk.f[j * n + a] = rho_b_in_d / 9;
                drop.f[j * n + a] = rho_droplet / 9;
            }
        }
    return 0;
}

}  // extern "C"
