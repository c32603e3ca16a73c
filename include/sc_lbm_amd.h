/*
 * sc_lbm_amd.h -- C-ABI of the MI355X (gfx950) D2Q9 Shan-Chen library
 * (libsclbm.so).  Plain pointers and sizes only; every call returns an int
 * status (SCL_OK = 0) and never throws; scl_last_error() returns the message
 * of the last failure on the calling thread.
 *
 * The reference (LBMFluids/sc_lbm) has no FFI: its operator API is the C++
 * classes LBM, Fluid and Geometry (include/lbm.h, include/fluid.h,
 * include/geometry.h of the reference).  The entry points below are what
 * those classes bind to; the drop-in classes in include/lbm.h, fluid.h and
 * geometry.h of this repository call exactly these.  Each declaration names
 * the reference method it replaces.
 *
 * Data model.
 *   scl_geometry  the solid mask of an nx x ny lattice, kept on the device as
 *                 one bit per site (1 = fluid), rows padded to 64 columns.
 *   scl_field     one fluid component on that lattice: populations (double
 *                 buffered, FP64, or FP32 / FP16 deviations f_i - w_i*rho0)
 *                 plus the reference's per-site arrays (density, velocity or
 *                 momentum, equilibrium velocity, forces, feq), all resident
 *                 in device memory.
 * Host arrays use the reference layout: site s = y*nx + x, population
 * f[i*nx*ny + s], directions rest, E, N, W, S, NE, NW, SW, SE
 * (reference include/fluid.h:24-39, include/lbm.h:126-128).
 */
#ifndef SC_LBM_AMD_H
#define SC_LBM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define SCL_OK 0
#define SCL_ERR_ARG 1        /* bad argument (sizes, null pointers, precision)   */
#define SCL_ERR_HIP 2        /* HIP runtime failure (no device, launch, memory)  */
#define SCL_ERR_RANGE 3      /* the reference would throw std::out_of_range      */
#define SCL_ERR_UNSUPPORTED 4/* operation not available for this precision       */
#define SCL_ERR_BUSY 5       /* server queue full: back-pressure, retry later     */
#define SCL_ERR_TIMEOUT 6    /* a wait ended before the job or server finished    */
#define SCL_ERR_INEXACT 7    /* SCL_FP64: a fused step reused an equilibrium velocity
                                older than the previous step's at a zero-density site (the
                                reference's rule, src/lbm.cpp:262-269, needs the previous
                                one): reported once, by the next download, moments or
                                checkpoint of the field; the data are still delivered.  A
                                served FP64 job fails with this reason. */

/* Population storage forms. */
#define SCL_FP64 0           /* plain FP64, bit-exact with the reference          */
#define SCL_FP32_DEV 1       /* FP32 deviations f_i - w_i*rho0, FP32 arithmetic   */
#define SCL_FP16_DEV 2       /* FP16 of (f_i - w_i*rho0)*2^8, FP32 arithmetic     */

/* Field selectors for scl_field_upload / scl_field_download. */
#define SCL_F 0              /* populations, 9*nx*ny                                 */
#define SCL_FEQ 1            /* equilibrium populations of the last collide, 9*nx*ny  */
#define SCL_RHO 2            /* density, nx*ny                                        */
#define SCL_UX 3             /* velocity (1 comp) or momentum (2 comp), nx*ny         */
#define SCL_UY 4
#define SCL_UEQX 5           /* equilibrium velocity, nx*ny                           */
#define SCL_UEQY 6
#define SCL_FRX 7            /* fluid-fluid (repulsive) force, nx*ny                  */
#define SCL_FRY 8
#define SCL_FSX 9            /* fluid-solid (adhesion) force, nx*ny                   */
#define SCL_FSY 10

typedef struct scl_geometry scl_geometry;
typedef struct scl_field scl_field;

/* Message of the last failed call on this thread ("" if none). */
const char* scl_last_error(void);

/* Select the HIP device used by the calling thread's subsequent creates
 * (default 0).  Fails if no gfx950 device is present. */
int scl_init(int device);

/* Block until all work queued by this library has finished. */
int scl_synchronize(void);

/* ---- Geometry: replaces Geometry's mask storage and accessors
 *      (include/geometry.h:58-59, 241-262 of the reference). ------------- */

/* mask: nx*ny ints, row-major, 1 = fluid, 0 = solid (anything else is
 * treated as solid except for the 1 test, as operator()==1 in the
 * reference's loops). */
int scl_geometry_create(int nx, int ny, const int* mask, scl_geometry** out);
int scl_geometry_destroy(scl_geometry* g);

/* ---- Field: replaces Fluid's storage, constructor and initialisers
 *      (include/fluid.h:60-70, src/fluid.cpp:43-151 of the reference). --- */

/* One component on geometry g.  omega is formed as 1/tau
 * (include/fluid.h:62).  precision is SCL_FP64 / SCL_FP32_DEV / SCL_FP16_DEV;
 * rho0 is the deviation reference density (ignored for SCL_FP64). All arrays
 * start at zero, as zero_density_ini (src/fluid.cpp:43-64). */
int scl_field_create(const scl_geometry* g, int precision, double tau, double rho0,
                     scl_field** out);
int scl_field_destroy(scl_field* f);

/* Fluid::initialize_interactions(Gs, Gf) (include/fluid.h:110-114). */
int scl_field_set_interactions(scl_field* f, double g_solid, double g_fluid);

/* Copy a field in / out; n must equal its size.  Populations are converted
 * to / from the storage form; the other arrays are FP64 and need SCL_FP64.
 * Uploading SCL_FSX / SCL_FSY sets the fluid-solid force from the host (the
 * reference's add_surface_forces); scl_step2 then refuses the field
 * (SCL_ERR_UNSUPPORTED) because the fused kernel derives that force from
 * the mask, and the per-operator entry points must be used. */
int scl_field_upload(scl_field* f, int which, const double* src, size_t n);
int scl_field_download(scl_field* f, int which, double* dst, size_t n);

/* Exact checkpoint of a field (no reference counterpart: the reference's
 * Fluid::restart_from_file re-reads 6-digit text, include/fluid.h:249-250,
 * src/fluid.cpp:285-300).  It holds the populations as stored (any form),
 * their phase and the equilibrium velocities of the zero-density rule, so a
 * field restored from it continues bit for bit as the original would.  The
 * fluid-solid force is not part of it: call scl_solid_surface_force on the
 * restored fields as after creation.  scl_field_restore needs a field of the
 * same size, precision and rho0 on a geometry with the same mask (the
 * checkpoint records a hash of it: solid sites hold bounce-back slots), and
 * returns SCL_ERR_ARG otherwise; bytes must equal
 * scl_field_checkpoint_size. */
int scl_field_checkpoint_size(const scl_field* f, size_t* bytes);
int scl_field_checkpoint(scl_field* f, void* dst, size_t bytes);
int scl_field_restore(scl_field* f, const void* src, size_t bytes);

/* ---- Per-operator entry points, reference semantics (SCL_FP64 only). --- */

/* Fluid::compute_density, src/fluid.cpp:175-183 */
int scl_compute_density(scl_field* f);
/* Fluid::compute_velocities, src/fluid.cpp:186-200 */
int scl_compute_velocities(const scl_geometry* g, scl_field* f);
/* Fluid::compute_f_equilibrium(geom), src/fluid.cpp:210-243 */
int scl_compute_f_equilibrium_geom(const scl_geometry* g, scl_field* f);
/* Fluid::compute_f_equilibrium(), src/fluid.cpp:246-278 */
int scl_compute_f_equilibrium(scl_field* f);
/* LBM::compute_solid_surface_force + Fluid::add_surface_forces,
 * src/lbm.cpp:102-141, src/fluid.cpp:154-160 (first call per field wins) */
int scl_solid_surface_force(const scl_geometry* g, scl_field* a, scl_field* b);
/* LBM::compute_fluid_repulsive_interactions, src/lbm.cpp:144-206 */
int scl_fluid_repulsion(const scl_geometry* g, scl_field* a, scl_field* b);
/* LBM::compute_equilibrium_velocities, src/lbm.cpp:209-272
 * (SCL_ERR_RANGE before scl_solid_surface_force, as the reference's .at()) */
int scl_equilibrium_velocities(const scl_geometry* g, scl_field* a, scl_field* b);
/* LBM::collide(geom, fluid), src/lbm.cpp:275-288 */
int scl_collide1(const scl_geometry* g, scl_field* f);
/* LBM::collide(fluid, fluid), src/lbm.cpp:291-312 */
int scl_collide2(scl_field* a, scl_field* b);
/* LBM::add_volume_force(geom, fluid, force), src/lbm.cpp:315-325 */
int scl_volume_force(const scl_geometry* g, scl_field* f, const double force[9]);
/* LBM::stream(geom, fluid), src/lbm.cpp:343-384 */
int scl_stream1(const scl_geometry* g, scl_field* f);
/* LBM::stream(geom, fluid, fluid), src/lbm.cpp:387-435 */
int scl_stream2(const scl_geometry* g, scl_field* a, scl_field* b);

/* ---- Fused steps: nsteps repetitions of the reference's step loop. ----- */

/* collide(geom,f); [add_volume_force(geom,f,force)]; stream(geom,f)
 * (benchmarks/single_phase_one_component_flows/laminar_flow_x_dir.cpp:64-68).
 * force may be NULL. */
int scl_step1(const scl_geometry* g, scl_field* f, const double* force, int nsteps);

/* a.compute_density(); b.compute_density();
 * compute_fluid_repulsive_interactions; compute_equilibrium_velocities;
 * collide(a,b); [add_volume_force(geom,a,force_a)];
 * [add_volume_force(geom,b,force_b)]; stream(geom,a,b)
 * (phase_separation.cpp:99-108, tests/lbm/lbm_tests.cpp:121-127).
 * force_a / force_b may be NULL.  Requires scl_solid_surface_force first,
 * as the reference does (SCL_ERR_RANGE otherwise). */
int scl_step2(const scl_geometry* g, scl_field* a, scl_field* b,
              const double* force_a, const double* force_b, int nsteps);

/* ---- Diagnostics (no reference counterpart: FP64 device reductions). -- */

/* Total mass and momentum of the current populations, summed in FP64 on the
 * device in a fixed order: out = {mass, momentum_x, momentum_y}. */
int scl_moments(const scl_geometry* g, scl_field* f, double out[3]);

/* scl_moments without allocating device memory (no reference counterpart),
 * for callers that cannot allocate: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Same result, bit for bit.
 * scl_moments stages the populations in an FP64 buffer the field allocates
 * on first use; this variant reads them as stored and keeps its partial sums
 * in a fixed device array of the library (3 doubles per row, at most 32768
 * rows: SCL_ERR_ARG beyond), one per device for the life of the process.
 * Calls that overlap on one device race for that array. */
int scl_moments_noalloc(const scl_geometry* g, scl_field* f, double out[3]);

/* ---- Droplet observables on the device.  The reference evaluates these in
 *      Python from the density files its drivers write; these entry points
 *      replace those scripts' arithmetic without downloading the fields.
 *      Densities are those of the fields' current streamed states (what
 *      Fluid::compute_density would give now), any storage form. ------- */

/* Laplace law, benchmarks/two_phase_two_components_flows/laplace_law/
 * compute_surface_tension.py:63-70 (G12, p_in, p_out, rho_bulk, rho_disp of
 * lines 34-43).  out[0] = p(x_in, y_in) - p(x_out, y_out) with
 * p = 1/3 (rho_b + rho_d) + G/3 rho_b rho_d; out[1] = the droplet radius
 * 0.5 (Rx + Ry), half the y and x extents of the sites with rho_d < rho_b,
 * rho_disp < rho_b <= rho_bulk.  SCL_ERR_RANGE for a site outside the
 * lattice or an empty set. */
int scl_laplace_observables(const scl_geometry* g, scl_field* droplet, scl_field* bulk, double G, int x_in,
                            int y_in, int x_out, int y_out, double rho_bulk, double rho_disp, double out[2]);

/* Contact angle of a droplet resting on a solid below it,
 * benchmarks/two_phase_two_components_flows/contact_angles/
 * compute_contact_angles.py:46-100: interface sites rho_b <= rho_d,
 * rho_d <= rho_bulk, rho_b > rho_disp; base from its two lowest rows, height
 * from its y extent.  out = {theta in degrees, base length L, height H,
 * radius R}.  SCL_ERR_RANGE if the interface is empty or has no site in the
 * row above its lowest. */
int scl_contact_angle(const scl_geometry* g, scl_field* droplet, scl_field* bulk, double rho_bulk, double rho_disp,
                      double out[4]);

/* Launch shape of the fused kernels for geometry g: columns per workgroup
 * (64, 128, 256 or 512) and rows per workgroup (>= 1).  0 = automatic.
 * Results do not depend on the shape (per-site arithmetic is fixed). */
int scl_set_tile(scl_geometry* g, int cols, int rows);
/* Rows of populations in flight in the fused kernel's software pipeline
 * (1 or 2; 0 = automatic: 1 for FP64, 2 for the deviation forms).  Results
 * do not depend on it. */
int scl_set_lookahead(scl_geometry* g, int rows_ahead);
/* Time steps per launch on g (no reference counterpart).  Two-component
 * fields on lattices at least 256 columns and 2K + 2 rows wide can advance
 * K = 2 (every form) or 3 (deviation forms) steps per launch (the
 * intermediate populations stay on chip; results are bit for bit those of
 * single steps).  0 (the default): what scl_autotune measured fastest, else
 * by lattice size: FP32 storage three per launch where two whole waves of
 * the three-step grid give bands of at least 16 rows, else two; FP16 two
 * where the two-step grid's bands reach 2 rows without solids or 10 with
 * them, else one; FP64 two; 1, 2
 * or 3: that many per launch wherever supported (a call's remaining steps go
 * two or one per launch; FP64 takes two while at least two remain, and a
 * read of the per-operator arrays after a call that ended in a two-step
 * launch recomputes that call's last step).  FP64 fields hand over to single steps
 * where a density falls below 1e-8 (the zero-density rule's held velocities
 * need them; the launch that met it is repeated).  rows: rows per band of
 * the multi-step kernels (0 = autotuned, else the height that makes the grid
 * two whole waves of resident workgroups; FP64 one wave where two would give
 * bands under 48 rows, and 16 rows above that on solid-free lattices; three
 * steps per launch with FP32 storage one wave on solid-free lattices). */
int scl_set_step_fusion(scl_geometry* g, int steps_per_launch, int rows);
/* Density below which FP64 steps on g hand over from two per launch to one
 * (no reference counterpart; default and least value 1e-8, the zero-density
 * rule's watch threshold).  A larger value only hands over earlier; results
 * do not depend on it. */
int scl_set_fp64_fusion_watch(scl_geometry* g, double rho);
/* What two-component steps of `precision` on g use: steps per launch (1, 2
 * or 3) and the band rows of the multi-step kernel (0 when one per
 * launch). */
int scl_get_step_fusion(const scl_geometry* g, int precision, int* steps_per_launch, int* rows);
/* The shape the fused kernel uses for `ncomp` components of `precision` on
 * g: the caller's overrides, else the autotuned one, else the default. */
int scl_get_tile(const scl_geometry* g, int ncomp, int precision, int* cols, int* rows, int* lookahead);

/* Autotuner (no reference counterpart): times the fused single step of
 * `ncomp` components of `precision` on g for every launch shape (columns
 * 64..512, rows 4..128, lookahead 1..2) on scratch fields at rest, keeps the
 * fastest for later steps on g and returns it (outputs may be NULL); for two
 * components in a deviation form it then times the two-step kernel over band
 * heights (16, 32, and those filling 1..4 whole waves of workgroups) and
 * keeps two steps per launch only if faster, then the three-step kernel
 * (bands of 32, 64 and 1..4 whole waves), kept only if faster than both.
 * Takes about 3 ms of GPU time per shape; needs memory for two scratch
 * fields. */
int scl_autotune(scl_geometry* g, int ncomp, int precision, int* cols, int* rows, int* lookahead);

/* Average duration in milliseconds of the fused kernel launches, measured
 * with HIP events on the library's stream.  reset > 0 clears the record and
 * starts recording, reset < 0 clears it and stops, reset == 0 reads it. */
int scl_timing(int reset, double* avg_ms, int* count);

/* Bandwidth probe (no reference counterpart): a plain device copy of
 * `bytes` bytes (kernel k_probe_copy, 16-byte vector loads and stores),
 * repeated `reps` times; avg_ms is the mean duration of one copy.  Used to
 * calibrate the HBM byte counters against a kernel of known traffic. */
int scl_probe_copy(size_t bytes, int reps, double* avg_ms);

/* ===========================================================================
 * Serving layer (no reference counterpart; the reference runs one lattice
 * per process).  A server keeps independent two-component Shan-Chen jobs
 * resident in device memory and advances all of them together: each step of
 * a round is one launch of a persistent kernel whose workgroups take tiles
 * (job, strip of columns, band of rows) from a work-list, and the steps of a
 * round are replayed from a captured HIP graph.  Requests are admitted
 * continuously, between rounds, as job slots and arena memory free up.
 * Per-site arithmetic is that of scl_step2 in the same precision, so a job's
 * result does not depend on the other jobs it shares the device with.
 * ======================================================================== */
typedef struct scl_server scl_server;

/* Generated initial conditions (densities; populations are rho/9 in every
 * direction on fluid sites and 0 on solid sites, as the reference's
 * initialisers, src/lbm.cpp:11-48).  Component B mirrors A (rho_hi <-> rho_lo). */
#define SCL_IC_UNIFORM 0     /* A = rho_hi everywhere                                   */
#define SCL_IC_DROPLET 1     /* A = rho_hi where (x-nx/2)^2+(y-ny/2)^2 <= radius^2      */
#define SCL_IC_STRATIFIED 2  /* A = rho_hi on rows y < ny/2                             */

/* Deliverables of a finished job (bit mask). */
#define SCL_DELIVER_RHO 1    /* final densities, nx*ny per component                  */
#define SCL_DELIVER_F 2      /* final populations (streamed state), 9*nx*ny per comp. */

/* Job states. */
#define SCL_JOB_QUEUED 0
#define SCL_JOB_RUNNING 1
#define SCL_JOB_DONE 2
#define SCL_JOB_FAILED 3     /* setup failed, or the mass watchdog tripped            */
#define SCL_JOB_CANCELLED 4
#define SCL_JOB_TIMED_OUT 5
#define SCL_JOB_SUSPENDED 6  /* pre-empted: exact state held in host memory           */

typedef struct scl_job {
    int nx, ny;                 /* lattice, 1..32767 each                             */
    /* geometry: caller's mask (nx*ny, 1 = fluid), or NULL for seeded disks */
    const int* mask;
    double porosity;            /* target fluid fraction of the generated mask (1 = open) */
    int rmin, rmax;             /* disk radii                                          */
    int walls;                  /* 1: solid rows y = 0 and y = ny-1                    */
    uint64_t geom_seed;
    /* initial state: caller's populations (9*nx*ny per component, reference
     * layout), or NULL for a generated initial condition */
    const double* f[2];
    int ic;                     /* SCL_IC_*                                            */
    double rho_hi, rho_lo, radius;
    double noise;               /* rho *= 1 + U[-noise, noise), seeded per site        */
    uint64_t ic_seed;           /* component A uses ic_seed, B ic_seed + 1             */
    /* physics (reference Fluid parameters, include/fluid.h:60-70,110-114) */
    double tau[2], g_fluid[2], g_solid[2];
    const double* force[2];     /* 9-vectors added on fluid sites each step, or NULL  */
    long long steps;            /* reference time steps to run, >= 1                   */
    double release_ms;          /* earliest admission, on the server clock            */
    double timeout_ms;          /* > 0: dropped (TIMED_OUT) if not done this long after release */
    int deliver;                /* SCL_DELIVER_* bits                                  */
    /* deviation forms: reference density rho0 per component (populations are
     * stored as f_i - w_i rho0); 0 = automatic: the midpoint of rho_hi and
     * rho_lo for a generated state, the mean density of the sites holding
     * populations for caller arrays; rounded to a multiple of 2^-10 */
    double rho0[2];
} scl_job;

typedef struct scl_job_status {
    int state;                  /* SCL_JOB_*                                           */
    long long steps_done;
    double release_ms, admit_ms, done_ms;   /* server clock; -1 if not reached        */
    double mass[2];             /* FP64 mass per component at the last round          */
    double mass0[2];            /* ... at the first round (watchdog reference)        */
    size_t bytes;               /* device memory held while resident                  */
    char error[160];
} scl_job_status;

typedef struct scl_server_config {
    int precision;              /* SCL_FP64, SCL_FP32_DEV or SCL_FP16_DEV              */
    size_t budget_bytes;        /* device arena for resident jobs (allocated once)    */
    int max_jobs;               /* resident job slots                                  */
    int max_pending;            /* queued requests before submit returns SCL_ERR_BUSY */
    int round_steps;            /* steps per round (graph length), 1..1024            */
    int tile_cols, tile_rows, lookahead;    /* persistent-kernel tile: columns 128 or 256,
                                   rows 1..4096 (0 = 16), lookahead 1 or 2; tile_cols 0 = the
                                   width is timed at start-up, and the lookahead too when 0
                                   (nonzero rows and lookahead are kept) */
    double mass_tol;            /* watchdog: |mass - mass0| > mass_tol*|mass0| fails a job;
                                   0 = by precision (1e-6 FP64, 1e-3 FP32, 0.1 FP16) */
    int max_bypass;             /* later jobs admitted past a waiting one before it blocks admission */
    int steps_per_launch;       /* deviation forms: jobs of at least 8 columns and 2K + 2 rows
                                   advance K = 2 or 3 steps per launch (same bits as 1): 1, 2
                                   (two), 3 (three, and two for jobs with two steps left in a
                                   round); 0 = automatic: 3 */
    int band_rows;              /* rows per band of the multi-step tiles; 0 = 32 (two-step),
                                   64 (three-step) */
    int use_graphs;             /* 1: each round replays a captured HIP graph; 0: the same
                                   launches issued one by one (profilers, debugging)      */
} scl_server_config;

typedef struct scl_server_stats {
    long long submitted, admitted, done, failed, cancelled, timed_out;
    long long rejected_busy;    /* submits refused by back-pressure                    */
    long long rejected;         /* submits refused as invalid or larger than the arena */
    long long rounds, job_steps, lattice_updates;
    double busy_ms;             /* wall time spent in rounds                           */
    double device_ms;           /* device time of the rounds' graphs (HIP events)     */
    double setup_ms;            /* wall time spent building admitted jobs' lattices    */
    double finish_ms;           /* wall time spent delivering finished jobs           */
    double clock_ms;            /* server clock now                                    */
    int resident, queued, max_resident;
    size_t arena_bytes, arena_used, arena_peak;
    int tile_cols, tile_rows, lookahead, round_steps, workgroups;
    int steps_per_launch;       /* the most steps per launch jobs take (1, 2 or 3)     */
    long long fluid_updates;    /* lattice updates of fluid sites only                 */
    long long pair_updates;     /* lattice updates made by multi-step launches         */
} scl_server_stats;

/* Defaults: FP32 deviations, 8 GiB arena, 64 slots, 4096 pending, rounds of
 * 36 steps (a multiple of the two- and three-step launches' 2 and 3, so full
 * rounds take no single steps), tile timed at start-up (16 rows), mass_tol by precision,
 * max_bypass 16, graphs on. */
int scl_server_default_config(scl_server_config* cfg);
/* Creates the server on the library's device and starts its request loop. */
int scl_server_create(const scl_server_config* cfg, scl_server** out);
/* Queues a job (copies everything it points to).  SCL_ERR_BUSY when
 * max_pending requests are queued; SCL_ERR_RANGE when the job can never fit
 * the arena; SCL_ERR_ARG for an invalid job. */
int scl_server_submit(scl_server* s, const scl_job* job, long long* id);
int scl_server_status(scl_server* s, long long id, scl_job_status* st);
/* Waits until the job has finished (any final state); SCL_ERR_TIMEOUT if
 * timeout_ms (< 0: forever) passes first. */
int scl_server_wait(scl_server* s, long long id, double timeout_ms, scl_job_status* st);
/* Queued jobs are dropped at once, resident ones at the next round boundary. */
int scl_server_cancel(scl_server* s, long long id);
/* Pre-emption (no reference counterpart): a queued job is set aside at once;
 * a resident one at the next round boundary, where its exact state (the
 * populations as stored and the zero-density rule's equilibrium velocities)
 * is copied to host memory and its device memory freed.  The job then waits
 * in SCL_JOB_SUSPENDED (scl_server_wait keeps waiting, scl_server_drain does
 * not count it) until scl_server_resume queues it again in its release
 * order; it continues bit for bit as if never interrupted.
 * Suspending a suspended job does nothing; SCL_ERR_ARG for a finished job
 * or, from resume, for a job that is not suspended (a resident job whose
 * suspension is still pending just keeps running). */
int scl_server_suspend(scl_server* s, long long id);
int scl_server_resume(scl_server* s, long long id);
/* Copies a deliverable of a finished job: which = SCL_RHO (nx*ny) or SCL_F
 * (9*nx*ny), component 0 or 1. */
int scl_server_result(scl_server* s, long long id, int which, int comp, double* out, size_t n);
/* Forgets a finished job and its deliverables. */
int scl_server_release(scl_server* s, long long id);
/* Waits until no job is queued or resident. */
int scl_server_drain(scl_server* s, double timeout_ms);
int scl_server_get_stats(scl_server* s, scl_server_stats* st);
/* Stops the loop (queued, resident and suspended jobs end CANCELLED) and
 * frees the arena. */
int scl_server_destroy(scl_server* s);

/* Host helper (no device needed): the mask a job would use, nx*ny ints. */
int scl_job_mask(const scl_job* job, int* mask, size_t n);

/* The geometry pipeline on the device.  Bit masks are row-major, ny rows of
 * ceil(nx/64) 64-bit words, bit x & 63 of word x >> 6 set where site x is
 * fluid (bits past nx zero): the layout the kernels read.
 *
 * scl_mask_raster replaces Geometry::add_walls followed by add_circle calls
 * (src/geometry.cpp:34-67, 132-175; the node sets of
 * src/geom_object/ellipse.cpp:12-33) on an all-fluid nx x ny lattice:
 * walls_along 0 none, 1 the reference's "x" (wall_thickness rows at the
 * bottom and the top), 2 its "y" (columns at the left and the right); then
 * n_circles triples (D, xc, yc).  Walls and circles are checked on the host
 * as the reference checks them (SCL_ERR_ARG with its message where it
 * throws; circles of diameter above 65535: SCL_ERR_RANGE), rasterised on the
 * device, copied to bits (n_words = ny * ceil(nx/64)) and their fluid sites
 * counted there.  device_ms (may be NULL): the device time of the raster and
 * count.  The circles may overlap. */
int scl_mask_raster(int nx, int ny, long long wall_thickness, int walls_along, const long long* circles,
                    int n_circles, uint64_t* bits, size_t n_words, long long* fluid_sites, double* device_ms);
/* A served job's mask through the server's own pipeline: the seeded disk
 * sampler on the host (centres and sizes only), the rasterising, packing and
 * fluid count on the device; SCL_ERR_HIP if the device count differs from
 * the sampler's.  The same bits as scl_job_mask packed.  A job with its own
 * mask gets that mask packed (no device work).  ms (may be NULL): [0] the
 * sampler's host time, [1] the device time. */
int scl_job_mask_bits(const scl_job* job, uint64_t* bits, size_t n_words, long long* fluid_sites, double* ms);

#ifdef __cplusplus
}
#endif
#endif
