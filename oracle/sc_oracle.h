/*
 * sc_oracle.h -- CPU oracle for the D2Q9 single/two-component Shan-Chen step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial restatement of the reference
 * algorithm (LBMFluids/sc_lbm, src/lbm.cpp and src/fluid.cpp) used by the
 * tests, by the smoke check and by the CPU-baseline leg of bench.py, always as
 * the checker and never as the thing measured or shipped.  The product path
 * (libsclbm.so) never links or calls it.
 *
 * Every function of the oracle is pinned by at least one of: the golden
 * vectors in tests/golden (written by the compiled reference), the reference's
 * own closed-form fixtures (tests/fluid/two_fluids_test_main.cpp), and the
 * mathematical invariants in tests/test_oracle.py.
 *
 * Layout conventions follow the reference (include/fluid.h:24-39):
 *   site index  s = y*nx + x
 *   population  f[i*N + s],  N = nx*ny,  i = 0..8 in the order
 *               rest, E, N, W, S, NE, NW, SW, SE
 *   mask        int per site, 1 = fluid, 0 = solid (include/geometry.h:15-33)
 */
#ifndef SC_ORACLE_H
#define SC_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sys oracle_sys;

/* Field selectors for oracle_get / oracle_set (per component). */
enum {
    ORACLE_F = 0,     /* populations, 9N                                  */
    ORACLE_FEQ = 1,   /* equilibrium populations of the last collide, 9N  */
    ORACLE_RHO = 2,   /* density, N                                       */
    ORACLE_UX = 3,    /* velocity (1 comp) or momentum (2 comp), N        */
    ORACLE_UY = 4,
    ORACLE_UEQX = 5,  /* equilibrium velocity (2 comp), N                 */
    ORACLE_UEQY = 6,
    ORACLE_FRX = 7,   /* fluid-fluid (repulsive) force (2 comp), N        */
    ORACLE_FRY = 8,
    ORACLE_FSX = 9,   /* fluid-solid (adhesion) force (2 comp), N         */
    ORACLE_FSY = 10
};

/* ncomp 1 or 2.  tau per component (omega = 1/tau is formed as in
 * include/fluid.h:62).  g_fluid / g_solid are the Shan-Chen couplings
 * (ignored for ncomp = 1).  Returns NULL on bad arguments. */
oracle_sys* oracle_create(int nx, int ny, int ncomp, const int* mask,
                          const double* tau, const double* g_fluid,
                          const double* g_solid);
void oracle_destroy(oracle_sys* s);

/* Copy a field in / out (n must equal the field size). Return 0 on success. */
int oracle_set(oracle_sys* s, int comp, int field, const double* src, size_t n);
int oracle_get(const oracle_sys* s, int comp, int field, double* dst, size_t n);

/* Once per run, two components: the fluid-solid force of
 * src/lbm.cpp:102-141 + src/fluid.cpp:154-160. */
void oracle_solid_surface_force(oracle_sys* s);

/* Canonical steps.  force[c] is a 9-vector or NULL (no volume force on that
 * component).  Two components follow
 *   density A, density B, repulsion, equilibrium velocities, collide,
 *   [volume force A], [volume force B], stream
 * (benchmarks/.../phase_separation.cpp:99-108, tests/lbm/lbm_tests.cpp:121-127);
 * one component follows collide, [volume force], stream
 * (benchmarks/.../laminar_flow_x_dir.cpp:64-68). */
void oracle_run(oracle_sys* s, int nsteps, const double* force0, const double* force1);

/* Individual operators (same semantics as the reference method cited). */
void oracle_density(oracle_sys* s, int comp);                    /* src/fluid.cpp:175-183 */
void oracle_velocities(oracle_sys* s, int comp);                 /* src/fluid.cpp:186-200 */
void oracle_feq_geom(oracle_sys* s, int comp);                   /* src/fluid.cpp:210-243 */
void oracle_feq_mc(oracle_sys* s, int comp);                     /* src/fluid.cpp:246-278 */
void oracle_repulsion(oracle_sys* s);                            /* src/lbm.cpp:144-206   */
void oracle_eq_velocities(oracle_sys* s);                        /* src/lbm.cpp:209-272   */
void oracle_collide1(oracle_sys* s);                             /* src/lbm.cpp:275-288   */
void oracle_collide2(oracle_sys* s);                             /* src/lbm.cpp:291-312   */
void oracle_volume_force(oracle_sys* s, int comp, const double* force); /* src/lbm.cpp:315-325 */
void oracle_stream(oracle_sys* s);                               /* src/lbm.cpp:343-435   */

/* Total mass and momentum per component, summed in long double, for the
 * conservation checks.  out = {mass, jx, jy} per component. */
void oracle_moments(const oracle_sys* s, int comp, double out[3]);

/* Geometry restatement: add_walls (src/geometry.cpp:34-69) and add_circle
 * (src/geometry.cpp:132-147 -> src/geom_object/ellipse.cpp:12-33, with the
 * bounds checks of src/geometry.cpp:150-175).  mask is nx*ny ints, updated in
 * place.  Return 0 on success, 1 on the reference's runtime_error, 2 on its
 * invalid_argument. */
int oracle_add_walls(int* mask, int nx, int ny, size_t dh, int where_x);
int oracle_add_circle(int* mask, int nx, int ny, size_t d, size_t xc, size_t yc);
int oracle_add_rectangle(int* mask, int nx, int ny, size_t lx, size_t ly, size_t xc, size_t yc);

/* Initialisers (src/lbm.cpp:11-48 and :51-99): write f of both components
 * of a two-component system in place. */
void oracle_init_droplet(oracle_sys* s, double rho_bulk, double rho_droplet,
                         double rho_b_in_d, double rho_d_in_b,
                         double xc, double yc, double radius);
int oracle_init_rectangle(oracle_sys* s, double rho_bulk, double rho_droplet,
                          double rho_b_in_d, double rho_d_in_b,
                          double xc, double yc, double half_lx, double half_ly);

#ifdef __cplusplus
}
#endif
#endif
