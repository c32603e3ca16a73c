// sclbm.hip -- runtime and C-ABI of libsclbm.so (see include/sc_lbm_amd.h).
//
// Owns the device state of geometries and fields, tracks the population
// phase of each field (reference streamed state S, or post-collision P whose
// streamed state is Gather(P)), dispatches the fused and per-operator
// kernels (two-component fields take several steps per launch: k_stepk,
// k_step2x and k_step2d, as the autotuner measured fastest or by lattice
// size), and recomputes the reference's intermediate arrays of the last
// fused step on demand ("replay": the input buffer of that step, or of the
// two-step launch that ended the call, is still intact, so rerunning from it
// in replay mode writes exactly the arrays the reference would hold).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <memory>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/sc_lbm_amd.h"
#include "../host/serve_core.h"
#include "lattice.h"
#include "ops.h"
#include "runtime.h"
#include "step.h"
#include "step2x.h"
#include "stepk.h"

using namespace scl;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(SCL_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                              \
    } while (0)

#define CHK(expr)                  \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != SCL_OK) return rc_; \
    } while (0)

// Device memory of the allocation-free entry point (scl_moments_noalloc):
// the per-row partial sums of the moment reduction and its result, for
// lattices of up to kScratchRows rows.  Each device holds its own copy until
// the process ends; every call overwrites what it reads, and calls that
// overlap on one device race for it.  Host code uses its device address
// (g_lib.scratch, from hipGetSymbolAddress), never the array's name.
constexpr int kScratchRows = 32768;
static __device__ __align__(256) double scratch[3 * kScratchRows + 3];

// ---------------------------------------------------------------------------
// library state
// ---------------------------------------------------------------------------
namespace {

struct Lib {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<int> ev_steps;   // time steps covered by each recorded launch
    int cus = 0;                 // compute units of the device
    int x2_occ[3][2] = {};       // [precision][solid-free]: two-step workgroups per CU (0 = not queried)
    double* scratch = nullptr;   // address of `scratch` on the library's device (set at initialisation)
    unsigned* f64_flag = nullptr;   // FP64 two-step launches: number of the first that met a dilute site
};
Lib g_lib;

const size_t kMaxTimed = 1 << 15;

int lib_ready() {
    if (g_lib.ready) return SCL_OK;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (n <= g_lib.device) return fail(SCL_ERR_HIP, "HIP device %d not present (%d devices)", g_lib.device, n);
    HIPCHK(hipSetDevice(g_lib.device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, g_lib.device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SCL_ERR_HIP, "device %d is %s; this library is built for gfx950", g_lib.device, prop.gcnArchName);
    g_lib.cus = prop.multiProcessorCount;
    // default priority (no priority argument): the library never changes it
    HIPCHK(hipStreamCreateWithFlags(&g_lib.stream, hipStreamNonBlocking));
    // the allocation-free entry point's array: its address on this device
    HIPCHK(hipGetSymbolAddress((void**)&g_lib.scratch, HIP_SYMBOL(scratch)));
    HIPCHK(hipMalloc((void**)&g_lib.f64_flag, 256));
    HIPCHK(hipMemset(g_lib.f64_flag, 0, 256));
    g_lib.ready = true;
    return SCL_OK;
}

template <class T>
int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return SCL_OK;
    HIPCHK(hipMalloc((void**)p, count * sizeof(T)));
    return SCL_OK;
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

}  // namespace

struct scl_geometry {
    int nx = 0, ny = 0, pitch = 0, wpr = 0;
    long long plane = 0;
    uint64_t* mask = nullptr;
    std::vector<int> host_mask;
    uint64_t mask_hash = 0;             // FNV-1a of the packed mask (checkpoints record it)
    // Launch shape of the fused kernel: the caller's choice (scl_set_tile /
    // scl_set_lookahead; 0 = not set), else the autotuner's per component
    // count and precision (scl_autotune), else the built-in default.
    int tile_cols = 0, tile_rows = 0;
    int lookahead = 0;
    int tuned[2][3][3] = {};            // [ncomp-1][precision] -> {cols, rows, lookahead}
    bool open = false;                  // no solid site: the fused kernel skips the mask
    int fusion = 0;                     // time steps per launch: 0 automatic, or 1..3 where supported
    int x2_rows = 0;                    // rows per band of the multi-step kernels (0 = automatic)
    int x2_tuned[3] = {};               // [precision]: autotuned band rows; -1: single steps were faster
    int x3_tuned[3] = {};               // [precision]: the same for three steps per launch; -1: slower
    double x2d_watch = kUeqWatch64;     // FP64 two-step launches hand over below this density
    double* part = nullptr;             // moment partials (3 per row + 3)
    int* obs = nullptr;                 // droplet-observable row extents (6 per row), allocated on first use
    // Fields keep their geometry alive: destroying a geometry that fields
    // still use only marks it, and the last field destroyed frees it.
    int refs = 0;
    bool dead = false;
    Grid2 grid() const { return Grid2{nx, ny, pitch, wpr, plane, mask}; }
};

namespace {
enum Phase { PH_S = 0, PH_P = 1 };

struct Replay;
}  // namespace

struct scl_field {
    scl_geometry* g = nullptr;
    int precision = SCL_FP64;
    double tau = 1.0, omega = 1.0, omc = 0.0, inv_omega = 1.0, rho0 = 0.0;
    double g_solid = 0.0, g_fluid = 0.0;
    bool has_solid_force = false;
    bool solid_force_uploaded = false;   // F_solid came from the host, not from the mask
    double gsneg_captured = 0.0;
    void* f[2] = {nullptr, nullptr};
    int cur = 0;
    int phase = PH_S;
    // The geometry whose mask produced the post-collision populations (phase
    // P).  Held by reference.
    scl_geometry* pg = nullptr;
    // Deviation forms: equilibrium velocities of the zero-density rule in
    // kUeqRing rotating buffer sets (x, y) of one block (UeqRing's layout);
    // a step reads set ucur and writes the next, so steps fused into one
    // launch (up to kUeqRing - 1 of them) never read what the same launch
    // writes.  ueqx / ueqy point at set ucur.
    double* ueq_block = nullptr;
    double* ueqs[kUeqRing][2] = {};
    int ucur = 0;
    // FP64: zero-density rule records (CompArgs::stamp), n_sites + 2 entries,
    // the last two the error flag and the step as of which every held
    // equilibrium velocity is valid; nstep = time steps completed (fused or
    // streamed per operator).
    unsigned* stamp = nullptr;
    unsigned nstep = 0;
    // FP64: a two-step launch met a site below the zero-density rule's watch
    // threshold; this field's steps go one per launch from then on
    bool dilute = false;
    // reference per-site arrays (FP64, pitched)
    double *rho = nullptr, *ux = nullptr, *uy = nullptr, *ueqx = nullptr, *ueqy = nullptr;
    double *frx = nullptr, *fry = nullptr, *fsx = nullptr, *fsy = nullptr, *feq = nullptr;
    double* stage = nullptr;   // FP64 staging for conversions
    std::shared_ptr<Replay> replay;
    size_t esize() const { return precision == SCL_FP64 ? 8 : (precision == SCL_FP32_DEV ? 4 : 2); }
};

namespace {

struct Replay {
    int nc = 0;
    bool gather = false;
    scl_field* fl[2] = {nullptr, nullptr};
    StepArgs<double> args{};
    // A call that ended in a two-step FP64 launch (P_{n+1} from P_{n-1}):
    // the input of its last step, P_n, was never stored.  `pre` recomputes
    // it from P_{n-1}, still intact in the launch's input buffer, into the
    // buffer of the launch's output; the replay (args) then reads it, and
    // `post` steps it again into the input buffer, which then holds P_{n+1}
    // bit for bit as the launch wrote it (cur_after: that buffer).
    bool two = false;
    StepArgs<double> pre{}, post{};
    int nw = 4, la = 1;
    int cur_after[2] = {0, 0};
};

void geometry_release(scl_geometry* g);

// Record that f's post-collision populations were produced with g's mask.
void set_pgeom(scl_field* f, scl_geometry* g) {
    if (f->pg == g) return;
    scl_geometry* old = f->pg;
    f->pg = g;
    g->refs++;
    if (old) geometry_release(old);
}

// The deviation forms' equilibrium-velocity buffers after `steps` steps.
void advance_ueq(scl_field* f, int steps) {
    f->nstep += (unsigned)steps;
    if (f->precision == SCL_FP64) return;
    f->ucur = (f->ucur + steps) % kUeqRing;
    f->ueqx = f->ueqs[f->ucur][0];
    f->ueqy = f->ueqs[f->ucur][1];
}

size_t n_sites(const scl_geometry* g) { return (size_t)g->pitch * g->ny; }
size_t n_pop(const scl_geometry* g) { return (size_t)g->plane * Q; }

// Declares every held equilibrium velocity of f valid as of step v (after an
// upload, a restore or the per-operator velocities, which set them all).
int set_ueq_valid(scl_field* f, unsigned v) {
    if (!f->stamp) return SCL_OK;   // FP64 fields keep the records
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, g_lib.stream, f->stamp + n_sites(f->g) + 1, v);
    HIPCHK(hipGetLastError());
    return SCL_OK;
}

// The zero-density rule's flag (ueq_reuse): SCL_ERR_INEXACT, once, if a
// fused step reused an equilibrium velocity older than the previous step's.
int check_ueq_flag(scl_field* f) {
    if (!f->stamp) return SCL_OK;
    unsigned flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, f->stamp + n_sites(f->g), sizeof flag, hipMemcpyDeviceToHost, g_lib.stream));
    HIPCHK(hipStreamSynchronize(g_lib.stream));
    if (!flag) return SCL_OK;
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, g_lib.stream, f->stamp + n_sites(f->g), 0u);
    HIPCHK(hipGetLastError());
    const double watch = kUeqWatch;
    return fail(SCL_ERR_INEXACT,
                "zero-density rule: a site's density fell from >= %g to <= 1e-16 within one step, and the fused "
                "kernel had not kept its previous equilibrium velocity (results may differ from the reference's "
                "at that site; the per-operator entry points keep every site's)",
                watch);
}

int ensure_aux(scl_field* f) {
    if (f->rho) return SCL_OK;
    const scl_geometry* g = f->g;
    const size_t n = n_sites(g);
    double** arrs[] = {&f->rho, &f->ux, &f->uy, &f->frx, &f->fry, &f->fsx, &f->fsy};
    for (double** a : arrs) {
        CHK(dalloc(a, n));
        HIPCHK(hipMemsetAsync(*a, 0, n * sizeof(double), g_lib.stream));
    }
    CHK(dalloc(&f->feq, n_pop(g)));
    HIPCHK(hipMemsetAsync(f->feq, 0, n_pop(g) * sizeof(double), g_lib.stream));
    return SCL_OK;
}

int ensure_stage(scl_field* f) {
    if (f->stage) return SCL_OK;
    return dalloc(&f->stage, n_pop(f->g));
}

dim3 site_grid(const scl_geometry* g) { return dim3((g->nx + 255) / 256, g->ny); }

// ---- fused-kernel dispatch ------------------------------------------------
struct Shape {
    int nw, rows, la;
};

// Default shape, from measurements on MI355X at 4096^2: 256-column strips
// (4 waves), about 2048 workgroups (8 per CU), one row in flight for the FP64
// and FP32 forms and two for the FP16 form (its register sets are smaller
// than its loads' latency budget).
Shape default_shape(const scl_geometry* g, int prec) {
    Shape s;
    s.nw = (g->nx + 63) / 64;
    if (s.nw > 4) s.nw = 4;
    if (s.nw == 3) s.nw = 4;
    const int nbx = (g->nx + 64 * s.nw - 1) / (64 * s.nw);
    long long r = (long long)nbx * g->ny / 2048;
    s.rows = (int)(r < 4 ? 4 : (r > 64 ? 64 : r));
    if (s.rows > g->ny) s.rows = g->ny;
    s.la = prec == SCL_FP16_DEV ? 2 : 1;
    return s;
}

Shape resolve_shape(const scl_geometry* g, int nc, int prec) {
    Shape s = default_shape(g, prec);
    const int* t = g->tuned[nc - 1][prec];
    if (t[0]) s.nw = t[0] / 64;
    if (t[1]) s.rows = t[1];
    if (t[2]) s.la = t[2];
    if (g->tile_cols) s.nw = g->tile_cols / 64;
    if (g->tile_rows) s.rows = g->tile_rows;
    if (g->lookahead) s.la = g->lookahead;
    if (s.rows > g->ny) s.rows = g->ny;
    return s;
}

int prec_of(size_t esize) { return esize == 8 ? SCL_FP64 : (esize == 4 ? SCL_FP32_DEV : SCL_FP16_DEV); }

// 32-bit byte offsets address one component's population buffer while it
// is below 4 GiB; larger buffers use the 64-bit variant (4-wave tiles).
bool needs_wide(const scl_geometry* g, size_t esize) {
    return (unsigned long long)g->plane * Q * esize >= (1ull << 32);
}

template <int NC, class ST, int NW, bool GATHER, bool REPLAY, class OffT, int LA, bool OPEN>
void launch_step_t(const StepArgs<typename ST::T>& a, int nblk) {
    hipLaunchKernelGGL((k_step<NC, ST, NW, GATHER, REPLAY, OffT, LA, OPEN>), dim3(nblk), dim3(64 * NW), 0,
                       g_lib.stream, a);
}

template <int NC, class ST, bool GATHER, int LA, bool OPEN>
int launch_nw_la(int nw, const StepArgs<typename ST::T>& a, int nblk) {
    switch (nw) {
        case 1: launch_step_t<NC, ST, 1, GATHER, false, uint32_t, LA, OPEN>(a, nblk); break;
        case 2: launch_step_t<NC, ST, 2, GATHER, false, uint32_t, LA, OPEN>(a, nblk); break;
        case 4: launch_step_t<NC, ST, 4, GATHER, false, uint32_t, LA, OPEN>(a, nblk); break;
        case 8: launch_step_t<NC, ST, 8, GATHER, false, uint32_t, LA, OPEN>(a, nblk); break;
        default: return fail(SCL_ERR_ARG, "unsupported tile width %d", 64 * nw);
    }
    return SCL_OK;
}

// The first step after an upload (no gather) runs once, so it has one
// shape per width; the gathering step also chooses the lookahead and the
// solid-free variant.
template <int NC, class ST, bool GATHER>
int launch_nw(int nw, int la, bool open, bool wide, const StepArgs<typename ST::T>& a, int nblk) {
    if (wide) {
        if (nw != 4) return fail(SCL_ERR_ARG, "lattices above 4 GiB per component use 256-column tiles");
        launch_step_t<NC, ST, 4, GATHER, false, uint64_t, 1, false>(a, nblk);
        return SCL_OK;
    }
    if (!GATHER) return launch_nw_la<NC, ST, GATHER, 1, false>(nw, a, nblk);
    if (open) {
        if (la == 2) return launch_nw_la<NC, ST, GATHER, 2, true>(nw, a, nblk);
        return launch_nw_la<NC, ST, GATHER, 1, true>(nw, a, nblk);
    }
    if (la == 2) return launch_nw_la<NC, ST, GATHER, 2, false>(nw, a, nblk);
    return launch_nw_la<NC, ST, GATHER, 1, false>(nw, a, nblk);
}

template <class T>
void fill_comp(CompArgs<T>& c, const scl_field* f, int src_buf, const double* force) {
    c.src = f->f[src_buf];
    c.dst = f->f[src_buf ^ 1];
    c.omega = (T)f->omega;
    c.omc = (T)f->omc;
    c.inv_omega = (T)f->inv_omega;
    c.gneg = (T)(-1.0 * f->g_fluid);
    c.gsneg = (T)f->gsneg_captured;
    c.rho0 = (T)f->rho0;
    set_force(c, force, f->omega);
    c.step = f->nstep + 1;
    c.stamp = f->stamp;
    if (f->precision == SCL_FP64) {
        c.ueqx = f->ueqx;
        c.ueqy = f->ueqy;
        c.ueqx_in = f->ueqx;
        c.ueqy_in = f->ueqy;
    } else {
        const int nxt = (f->ucur + 1) % kUeqRing;
        c.ueqx_in = f->ueqs[f->ucur][0];
        c.ueqy_in = f->ueqs[f->ucur][1];
        c.ueqx = f->ueqs[nxt][0];
        c.ueqy = f->ueqs[nxt][1];
    }
    c.rho = f->rho;
    c.jx = f->ux;
    c.jy = f->uy;
    c.frx = f->frx;
    c.fry = f->fry;
    c.feq = f->feq;
}

template <class T>
StepArgs<T> make_args(const scl_geometry* g, int rows, int nw) {
    StepArgs<T> a{};
    a.nx = g->nx;
    a.ny = g->ny;
    a.pitch = g->pitch;
    a.wpr = g->wpr;
    a.plane = g->plane;
    a.mask = g->mask;
    a.rows = rows;
    a.gx = (g->nx + 64 * nw - 1) / (64 * nw);
    return a;
}

int nblocks(const scl_geometry* g, int nw, int rows) {
    const int gx = (g->nx + 64 * nw - 1) / (64 * nw);
    const int gy = (g->ny + rows - 1) / rows;
    return gx * gy;
}

template <int NC, class ST>
int launch_fused(const scl_geometry* g, scl_field* const* fl, const double* const* forces, bool gather,
                 std::shared_ptr<Replay>* rec) {
    using T = typename ST::T;
    const Shape sh = resolve_shape(g, NC, prec_of(sizeof(typename ST::S)));
    int nw = sh.nw, rows = sh.rows;
    StepArgs<T> a = make_args<T>(g, rows, nw);
    for (int c = 0; c < NC; ++c) fill_comp(a.c[c], fl[c], fl[c]->cur, forces[c]);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (g_lib.timing && g_lib.ev.size() < kMaxTimed) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, g_lib.stream));
    }
    const bool wide = needs_wide(g, sizeof(typename ST::S));
    if (wide && nw != 4) {
        nw = 4;
        a = make_args<T>(g, rows, nw);
        for (int c = 0; c < NC; ++c) fill_comp(a.c[c], fl[c], fl[c]->cur, forces[c]);
    }
    const int nblk = nblocks(g, nw, rows);
    const int la = sh.la;
    int rc = gather ? launch_nw<NC, ST, true>(nw, la, g->open, wide, a, nblk)
                    : launch_nw<NC, ST, false>(nw, la, g->open, wide, a, nblk);
    if (rc != SCL_OK) return rc;
    HIPCHK(hipGetLastError());
    if (e0) {
        HIPCHK(hipEventRecord(e1, g_lib.stream));
        g_lib.ev.emplace_back(e0, e1);
        g_lib.ev_steps.push_back(1);
    }
    if (rec) {
        if constexpr (std::is_same<ST, StoreF64>::value) {
            auto r = std::make_shared<Replay>();
            r->nc = NC;
            r->gather = gather;
            for (int c = 0; c < NC; ++c) r->fl[c] = fl[c];
            // the replay reads the same source buffers; its outputs are the
            // reference arrays, so the replay launch uses the 4-wave shape
            StepArgs<double> ra = make_args<double>(g, rows, 4);
            for (int c = 0; c < NC; ++c) ra.c[c] = a.c[c];
            r->args = ra;
            *rec = r;
        }
    }
    return SCL_OK;
}

// Two fused steps (deviation forms, two components, gathering): P_{n+1}
// from P_{n-1} in one launch (k_step2x).
bool x2_supported(const scl_geometry* g, int precision) {
    static const size_t esz[3] = {8, 4, 2};
    return precision != SCL_FP64 && g->nx >= kX2Threads && g->ny >= xk_min_rows<2>() &&
           !needs_wide(g, esz[precision]);
}

int x2_bands(const scl_geometry* g, int precision, int waves);

// whether two-component steps of `precision` on g go two per launch:
// fusion >= 2, or what scl_autotune measured fastest; untuned, always with
// FP32 storage, and with FP16 storage where the lattice gives the two-step
// grid's two whole waves bands of at least 2 rows (solid-free; 10 with
// solids) before the 8-row minimum.  Below that the FP16 single step is
// faster: two-component MLUPS, one step against two per launch, 512^2 open
// 24.6k / 21.2k, 1024^2 with 20% solid disks 25.1k / 21.6k, 1024 x 512
// porous 25.9k / 18.5k; above it two per launch win: 1024^2 open 38.5k /
// 52.8k, 2048^2 with disks 30.6k / 37.4k.  (FP32 storage: two per launch
// won on every lattice measured, 512^2 open 20.7k / 22.1k.)
bool x2_wanted(const scl_geometry* g, int precision) {
    if (g->fusion == 1 || !x2_supported(g, precision)) return false;
    if (g->fusion >= 2 || g->x2_tuned[precision] > 0) return true;
    if (g->x2_tuned[precision] < 0) return false;
    if (precision != SCL_FP16_DEV) return true;
    return (double)g->ny / x2_bands(g, precision, 2) >= (g->open ? 2.0 : 10.0);
}

// Workgroups of the two-step kernel resident per CU (registers and LDS), of
// its force-free variant: the one the autotuner times (FP16 solid-free: four,
// the variant with a force three).
int x2_occupancy(int precision, bool open) {
    int& o = g_lib.x2_occ[precision][open ? 1 : 0];
    if (!o) {
        const void* k = precision == SCL_FP64
                            ? (open ? reinterpret_cast<const void*>(&k_step2d<true, false>)
                                    : reinterpret_cast<const void*>(&k_step2d<false, false>))
                        : precision == SCL_FP32_DEV
                            ? (open ? reinterpret_cast<const void*>(&k_step2x<StoreF32, true, false>)
                                    : reinterpret_cast<const void*>(&k_step2x<StoreF32, false, false>))
                            : (open ? reinterpret_cast<const void*>(&k_step2x<StoreF16, true, false>)
                                    : reinterpret_cast<const void*>(&k_step2x<StoreF16, false, false>));
        const int threads = precision != SCL_FP64 ? kX2Threads : (open ? x2d_threads<true>() : x2d_threads<false>());
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, threads, 0) != hipSuccess || n < 1) n = 1;
        o = n;
    }
    return o;
}

// Bands that make the two-step grid `waves` whole waves of resident
// workgroups (a partial last wave leaves CUs idle at the end of every
// launch), and their height (at least 8 rows).
int x2_bands(const scl_geometry* g, int precision, int waves) {
    const int out = precision == SCL_FP64       ? (g->open ? xk_out<StoreF64, 2, x2d_threads<true>()>()
                                                           : xk_out<StoreF64, 2, x2d_threads<false>()>())
                    : precision == SCL_FP32_DEV ? x2_out<StoreF32>()
                                                : x2_out<StoreF16>();
    const int strips = (g->nx + out - 1) / out;
    const int slots = x2_occupancy(precision, g->open) * std::max(g_lib.cus, 1);
    return std::max(1, waves * slots / strips);
}

int x2_wave_rows(const scl_geometry* g, int precision, int waves) {
    const int bands = x2_bands(g, precision, waves);
    return std::max(8, std::min((g->ny + bands - 1) / bands, g->ny));
}

// Default rows per band: two whole waves (52 rows at 4096^2 with FP32
// storage, 35 with FP16).  FP64 (one workgroup per CU with solids, two
// without) takes one wave where two would give bands under 48 rows: two-step
// MLUPS at one against two waves, 1024 x 512 porous 15.7k (11 rows) / 10.4k
// (8), 2048^2 open 27.6k (49) / 26.4k (25), 2048^2 with disks 23.6k (74) /
// 22.5k (37).  Above that, solid-free FP64 lattices take bands of 16 rows
// (many waves of short bands): 4096^2 28.2k against 26.3k at two waves (94
// rows) and 25.8k at one (188); 3072^2 28.0k (16) / 28.5k (70) / 24.9k
// (140).  With solids two waves stay (4096^2 with disks: 23.2k at 137 rows
// against 20.6k at 16).
int x2_rows_for(const scl_geometry* g, int precision) {
    int r = g->x2_rows;
    if (!r && g->x2_tuned[precision] > 0) r = g->x2_tuned[precision];
    if (!r) {
        r = x2_wave_rows(g, precision, 2);
        if (precision == SCL_FP64) r = r < 48 ? x2_wave_rows(g, precision, 1) : (g->open ? 16 : r);
    }
    return std::min(r, g->ny);
}

bool x2_ok(const scl_geometry* g, scl_field* const* fl, int nc) {
    return nc == 2 && fl[0]->phase == PH_P && fl[1]->phase == PH_P && x2_wanted(g, fl[0]->precision);
}

// Arguments of a multi-step launch (k_step2x, k_stepk): parameters and
// P_{n-1} of both fields, P_{n+K-1} into the other buffer, the
// equilibrium-velocity sets from the fields' current ones.
template <int K>
XkArgs<K> xk_args(const scl_geometry* g, scl_field* const* fl, const double* const* forces, int rows, int out) {
    XkArgs<K> A{};
    A.a = make_args<float>(g, rows, 4);
    A.a.gx = (g->nx + out - 1) / out;
    for (int c = 0; c < 2; ++c) {
        const scl_field* f = fl[c];
        fill_comp(A.a.c[c], f, f->cur, forces[c]);
        A.dst[c] = f->f[f->cur ^ 1];
        A.u.base[c] = f->ueq_block;
        A.u.cur[c] = f->ucur;
    }
    A.u.n = (long long)n_sites(g);
    A.u.set_stride = 2 * A.u.n;
    A.u.nsets = kUeqRing;
    return A;
}

template <class ST>
int launch_x2(const scl_geometry* g, scl_field* const* fl, const double* const* forces) {
    const int rows = x2_rows_for(g, fl[0]->precision);
    const XkArgs<2> A = xk_args<2>(g, fl, forces, rows, x2_out<ST>());
    const int nblk = A.a.gx * ((g->ny + rows - 1) / rows);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (g_lib.timing && g_lib.ev.size() < kMaxTimed) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, g_lib.stream));
    }
    // kernels without the force terms where neither component has a force
    const bool force = forces[0] || forces[1];
    const dim3 grid(nblk), block(kX2Threads);
    if (g->open) {
        if (force) hipLaunchKernelGGL((k_step2x<ST, true, true>), grid, block, 0, g_lib.stream, A);
        else hipLaunchKernelGGL((k_step2x<ST, true, false>), grid, block, 0, g_lib.stream, A);
    } else {
        if (force) hipLaunchKernelGGL((k_step2x<ST, false, true>), grid, block, 0, g_lib.stream, A);
        else hipLaunchKernelGGL((k_step2x<ST, false, false>), grid, block, 0, g_lib.stream, A);
    }
    HIPCHK(hipGetLastError());
    if (e0) {
        HIPCHK(hipEventRecord(e1, g_lib.stream));
        g_lib.ev.emplace_back(e0, e1);
        g_lib.ev_steps.push_back(2);
    }
    return SCL_OK;
}

// Three fused steps (k_stepk with K = 3): P_{n+2} from P_{n-1} in one
// launch.  Automatic choice (fusion 0): what scl_autotune measured fastest,
// else three per launch with FP32 storage (72k against 61k MLUPS for two on
// a 4096^2 lattice) where two whole waves of the three-step grid give bands
// of at least 16 rows (2048^2 open: 72.5k against 52.9k for two; 2048^2
// with disks 47.9k / 43.0k), two below (1024^2 open 43.5k / 51.2k; 1024 x
// 512 porous 20.5k / 24.2k), and two with FP16 storage (74k against 82k).
int x3_wave_rows(const scl_geometry* g, int precision, int waves);

bool x3_wanted(const scl_geometry* g, int precision) {
    if (!x2_supported(g, precision) || g->ny < xk_min_rows<3>() || g->fusion == 1 || g->fusion == 2) return false;
    if (g->fusion == 3) return true;
    if (g->x3_tuned[precision]) return g->x3_tuned[precision] > 0;
    return precision == SCL_FP32_DEV && x3_wave_rows(g, precision, 2) >= 16;
}

template <class ST, bool OPEN>
int x3_occupancy() {
    static int o = 0;
    if (!o) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void*>(&k_stepk<ST, OPEN, 3, x3_threads<ST, OPEN>(), false>),
                x3_threads<ST, OPEN>(), 0) != hipSuccess ||
            n < 1)
            n = 1;
        o = n;
    }
    return o;
}

// Band height that makes the three-step grid `waves` whole waves of
// resident workgroups (78 rows at 4096^2 with FP32 storage for two waves).
template <class ST, bool OPEN>
int x3_wave_rows_t(const scl_geometry* g, int waves) {
    constexpr int out = xk_out<ST, 3, x3_threads<ST, OPEN>()>();
    const int strips = (g->nx + out - 1) / out;
    const int bands = std::max(1, waves * x3_occupancy<ST, OPEN>() * std::max(g_lib.cus, 1) / strips);
    return std::max(8, std::min((g->ny + bands - 1) / bands, g->ny));
}

int x3_wave_rows(const scl_geometry* g, int precision, int waves) {
    if (precision == SCL_FP32_DEV)
        return g->open ? x3_wave_rows_t<StoreF32, true>(g, waves) : x3_wave_rows_t<StoreF32, false>(g, waves);
    return g->open ? x3_wave_rows_t<StoreF16, true>(g, waves) : x3_wave_rows_t<StoreF16, false>(g, waves);
}

// rows per band: the caller's, else the autotuned, else two whole waves, or
// one with FP32 storage on a solid-free lattice: three-step MLUPS at one
// against two waves, 4096^2 94.4k (158 rows) / 90.6k (78), 2048^2 78.7k
// (42) / 69.2k (21); with solids two stay (4096^2 with disks 57.2k at 90
// rows against 36.3k at 178).
int x3_rows_for(const scl_geometry* g, int precision) {
    const int waves = precision == SCL_FP32_DEV && g->open ? 1 : 2;
    const int r = g->x2_rows ? g->x2_rows
                             : (g->x3_tuned[precision] > 0 ? g->x3_tuned[precision] : x3_wave_rows(g, precision, waves));
    return std::min(r, g->ny);
}

template <class ST, bool OPEN>
int launch_x3_t(const scl_geometry* g, scl_field* const* fl, const double* const* forces) {
    constexpr int T = x3_threads<ST, OPEN>();
    constexpr int OUT = xk_out<ST, 3, T>();
    const int rows = x3_rows_for(g, fl[0]->precision);
    const XkArgs<3> A = xk_args<3>(g, fl, forces, rows, OUT);
    const int nblk = A.a.gx * ((g->ny + rows - 1) / rows);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (g_lib.timing && g_lib.ev.size() < kMaxTimed) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, g_lib.stream));
    }
    if (forces[0] || forces[1])
        hipLaunchKernelGGL((k_stepk<ST, OPEN, 3, T, true>), dim3(nblk), dim3(T), 0, g_lib.stream, A);
    else
        hipLaunchKernelGGL((k_stepk<ST, OPEN, 3, T, false>), dim3(nblk), dim3(T), 0, g_lib.stream, A);
    HIPCHK(hipGetLastError());
    if (e0) {
        HIPCHK(hipEventRecord(e1, g_lib.stream));
        g_lib.ev.emplace_back(e0, e1);
        g_lib.ev_steps.push_back(3);
    }
    return SCL_OK;
}

template <class ST>
int launch_x3(const scl_geometry* g, scl_field* const* fl, const double* const* forces) {
    return g->open ? launch_x3_t<ST, true>(g, fl, forces) : launch_x3_t<ST, false>(g, fl, forces);
}

// Two fused FP64 steps (k_step2d): P_{n+1} from P_{n-1} with the reference's
// arithmetic, bit for bit two single steps.  A site below the zero-density
// rule's watch threshold (whose held equilibrium velocity lives in HBM, with
// the records the single-step kernel keeps) stops the sequence: the launch
// writes its number into g_lib.f64_flag, later launches return at once, and
// fused_steps repeats the steps from that launch's input one per launch.
bool x2d_supported(const scl_geometry* g) {
    return g->nx >= (g->open ? x2d_threads<true>() : x2d_threads<false>()) && g->ny >= xk_min_rows<2>() &&
           !needs_wide(g, sizeof(double));
}

bool x2d_wanted(const scl_geometry* g) {
    if (g->fusion == 1 || !x2d_supported(g)) return false;
    return g->fusion >= 2 || g->x2_tuned[SCL_FP64] >= 0;
}

bool x2d_ok(const scl_geometry* g, scl_field* const* fl, int nc) {
    return nc == 2 && fl[0]->precision == SCL_FP64 && fl[0]->phase == PH_P && fl[1]->phase == PH_P &&
           !fl[0]->dilute && !fl[1]->dilute && x2d_wanted(g);
}

// The replay record of a call whose last launch is the two-step one about
// to run on fl (their cur and nstep are still the launch's input's): the
// single steps n and n + 1 with the shape launch_fused gives them.
std::shared_ptr<Replay> x2d_replay(const scl_geometry* g, scl_field* const* fl, const double* const* forces) {
    const Shape sh = resolve_shape(g, 2, SCL_FP64);
    auto r = std::make_shared<Replay>();
    r->nc = 2;
    r->gather = true;
    r->two = true;
    r->nw = sh.nw;
    r->la = sh.la;
    r->pre = make_args<double>(g, sh.rows, sh.nw);
    r->post = make_args<double>(g, sh.rows, sh.nw);
    for (int c = 0; c < 2; ++c) {
        r->fl[c] = fl[c];
        r->cur_after[c] = fl[c]->cur;
        fill_comp(r->pre.c[c], fl[c], fl[c]->cur, forces[c]);        // P_{n-1} -> P_n
        fill_comp(r->post.c[c], fl[c], fl[c]->cur ^ 1, forces[c]);   // P_n -> P_{n+1}
        r->post.c[c].step += 1;
    }
    r->args = make_args<double>(g, sh.rows, 4);
    for (int c = 0; c < 2; ++c) r->args.c[c] = r->post.c[c];
    return r;
}

int launch_x2d(const scl_geometry* g, scl_field* const* fl, const double* const* forces, unsigned launch) {
    const int T = g->open ? x2d_threads<true>() : x2d_threads<false>();
    const int OUT = g->open ? xk_out<StoreF64, 2, x2d_threads<true>()>() : xk_out<StoreF64, 2, x2d_threads<false>()>();
    const int rows = x2_rows_for(g, SCL_FP64);
    XkArgs<2, double> A{};
    A.a = make_args<double>(g, rows, 4);
    A.a.gx = (g->nx + OUT - 1) / OUT;
    for (int c = 0; c < 2; ++c) {
        fill_comp(A.a.c[c], fl[c], fl[c]->cur, forces[c]);
        A.dst[c] = fl[c]->f[fl[c]->cur ^ 1];
    }
    A.flag = g_lib.f64_flag;
    A.launch = launch;
    A.watch = g->x2d_watch;
    const int nblk = A.a.gx * ((g->ny + rows - 1) / rows);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (g_lib.timing && g_lib.ev.size() < kMaxTimed) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, g_lib.stream));
    }
    const bool force = forces[0] || forces[1];
    const dim3 grid(nblk), block(T);
    if (g->open) {
        if (force) hipLaunchKernelGGL((k_step2d<true, true>), grid, block, 0, g_lib.stream, A);
        else hipLaunchKernelGGL((k_step2d<true, false>), grid, block, 0, g_lib.stream, A);
    } else {
        if (force) hipLaunchKernelGGL((k_step2d<false, true>), grid, block, 0, g_lib.stream, A);
        else hipLaunchKernelGGL((k_step2d<false, false>), grid, block, 0, g_lib.stream, A);
    }
    HIPCHK(hipGetLastError());
    if (e0) {
        HIPCHK(hipEventRecord(e1, g_lib.stream));
        g_lib.ev.emplace_back(e0, e1);
        g_lib.ev_steps.push_back(2);
    }
    return SCL_OK;
}

// The number of the first FP64 two-step launch since the last call that met
// a dilute site (0: none); clears it.
int x2d_failed(unsigned* launch) {
    unsigned v = 0;
    HIPCHK(hipMemcpyAsync(&v, g_lib.f64_flag, sizeof v, hipMemcpyDeviceToHost, g_lib.stream));
    HIPCHK(hipStreamSynchronize(g_lib.stream));
    if (v) HIPCHK(hipMemsetAsync(g_lib.f64_flag, 0, sizeof v, g_lib.stream));
    *launch = v;
    return SCL_OK;
}

template <int NC, bool GATHER>
void replay_t(const Replay& r, int nblk, bool wide) {
    if (wide)
        hipLaunchKernelGGL((k_step<NC, StoreF64, 4, GATHER, true, uint64_t, 1, false>), dim3(nblk), dim3(256), 0,
                           g_lib.stream, r.args);
    else
        hipLaunchKernelGGL((k_step<NC, StoreF64, 4, GATHER, true, uint32_t, 1, false>), dim3(nblk), dim3(256), 0,
                           g_lib.stream, r.args);
}

int run_replay(Replay& r) {
    const scl_geometry* g = r.fl[0]->g;
    const int nblk = nblocks(g, 4, r.args.rows);
    const bool wide = needs_wide(g, sizeof(double));
    if (r.two) {   // two-component, gathering, not wide (two-step launches)
        CHK((launch_nw<2, StoreF64, true>(r.nw, r.la, g->open, false, r.pre, nblocks(g, r.nw, r.pre.rows))));
        HIPCHK(hipGetLastError());
        replay_t<2, true>(r, nblk, false);
        HIPCHK(hipGetLastError());
        CHK((launch_nw<2, StoreF64, true>(r.nw, r.la, g->open, false, r.post, nblocks(g, r.nw, r.post.rows))));
        HIPCHK(hipGetLastError());
        for (int c = 0; c < 2; ++c)
            if (r.fl[c]) r.fl[c]->cur = r.cur_after[c];
        return SCL_OK;
    }
    if (r.nc == 1) {
        if (r.gather) replay_t<1, true>(r, nblk, wide);
        else replay_t<1, false>(r, nblk, wide);
    } else {
        if (r.gather) replay_t<2, true>(r, nblk, wide);
        else replay_t<2, false>(r, nblk, wide);
    }
    HIPCHK(hipGetLastError());
    return SCL_OK;
}

// Write the intermediate arrays of the last fused step touching f, if any.
int materialize(scl_field* f) {
    if (!f->replay) return SCL_OK;
    std::shared_ptr<Replay> r = f->replay;
    CHK(run_replay(*r));
    for (int c = 0; c < r->nc; ++c)
        if (r->fl[c] && r->fl[c]->replay == r) r->fl[c]->replay.reset();
    return SCL_OK;
}

// Bring the populations of f to the reference's streamed state in place.
int ensure_s(scl_field* f) {
    if (f->phase == PH_S) return SCL_OK;
    CHK(materialize(f));
    const Grid2 gr = (f->pg ? f->pg : f->g)->grid();
    const int nxt = f->cur ^ 1;
    switch (f->precision) {
        case SCL_FP64:
            hipLaunchKernelGGL(k_gather<StoreF64>, site_grid(f->g), dim3(256), 0, g_lib.stream, gr,
                               (const double*)f->f[f->cur], (double*)f->f[nxt]);
            break;
        case SCL_FP32_DEV:
            hipLaunchKernelGGL(k_gather<StoreF32>, site_grid(f->g), dim3(256), 0, g_lib.stream, gr,
                               (const float*)f->f[f->cur], (float*)f->f[nxt]);
            break;
        default:
            hipLaunchKernelGGL(k_gather<StoreF16>, site_grid(f->g), dim3(256), 0, g_lib.stream, gr,
                               (const _Float16*)f->f[f->cur], (_Float16*)f->f[nxt]);
            break;
    }
    HIPCHK(hipGetLastError());
    f->cur = nxt;
    f->phase = PH_S;
    return SCL_OK;
}

// Every per-operator entry: keep the last step's intermediates, then work
// on the streamed state.
int prepare_op(scl_field* f) {
    if (!f) return fail(SCL_ERR_ARG, "null field");
    if (f->precision != SCL_FP64)
        return fail(SCL_ERR_UNSUPPORTED, "per-operator calls need an SCL_FP64 field");
    CHK(ensure_aux(f));
    CHK(materialize(f));
    CHK(ensure_s(f));
    return SCL_OK;
}

int check_geom(const scl_geometry* g, const scl_field* f) {
    if (!g || !f) return fail(SCL_ERR_ARG, "null geometry or field");
    if (g->nx != f->g->nx || g->ny != f->g->ny)
        return fail(SCL_ERR_ARG, "geometry %dx%d does not match the field's %dx%d", g->nx, g->ny, f->g->nx,
                    f->g->ny);
    return SCL_OK;
}

// Drop or materialize replay records before a fused step on `fl`.
int settle_replays(scl_field* const* fl, int nc) {
    for (int c = 0; c < nc; ++c) {
        std::shared_ptr<Replay> r = fl[c]->replay;
        if (!r) continue;
        bool same = r->nc == nc;
        for (int k = 0; k < r->nc && same; ++k) same = (r->fl[k] == fl[0] || (nc == 2 && r->fl[k] == fl[1]));
        if (same) {
            for (int k = 0; k < r->nc; ++k)
                if (r->fl[k]) r->fl[k]->replay.reset();
        } else {
            CHK(materialize(fl[c]));
        }
    }
    return SCL_OK;
}

}  // namespace

namespace {
template <int NC>
int fused_steps(const scl_geometry* g, scl_field** fl, const double* const* forces, int nsteps) {
    const int prec = fl[0]->precision;
    for (int c = 0; c < NC; ++c) {
        CHK(check_geom(g, fl[c]));
        if (fl[c]->precision != prec) return fail(SCL_ERR_ARG, "components have different precisions");
        if (prec == SCL_FP64) CHK(ensure_aux(fl[c]));
    }
    if (NC == 2) {
        if (fl[0] == fl[1]) return fail(SCL_ERR_ARG, "the two components must be different fields");
        if (!fl[0]->has_solid_force || !fl[1]->has_solid_force)
            return fail(SCL_ERR_RANGE, "fluid-solid force not computed (compute_solid_surface_force first)");
        if (fl[0]->solid_force_uploaded || fl[1]->solid_force_uploaded)
            return fail(SCL_ERR_UNSUPPORTED,
                        "fluid-solid force was uploaded from the host; the fused step derives it from the mask "
                        "(use the per-operator entry points)");
    }
    if (nsteps <= 0) return SCL_OK;
    CHK(settle_replays(fl, NC));
    // post-collision populations from another mask: back to the streamed
    // state with that mask, then this step starts from it
    for (int c = 0; c < NC; ++c)
        if (fl[c]->phase == PH_P && fl[c]->pg != g) CHK(ensure_s(fl[c]));
    // both components must be in the same phase
    if (NC == 2 && fl[0]->phase != fl[1]->phase) {
        CHK(ensure_s(fl[0]));
        CHK(ensure_s(fl[1]));
    }
    // FP64 two-step launches go in runs that end in a check of the flag
    // (x2d_failed): the run's first step, buffers and step counts, to go
    // back to the input of a launch that met a dilute site
    unsigned run = 0;
    int run_t = 0, run_cur[NC];
    unsigned run_nstep[NC];
    std::shared_ptr<Replay> last2;   // the replay record if the call ends in a two-step launch
    for (int t = 0; t < nsteps || run;) {
        if constexpr (NC == 2) {
            // FP64: two steps per launch while at least two remain; a call
            // that ends in one keeps a replay record that recomputes its
            // last step's input (x2d_replay), one that ends in a single step
            // that step's (launch_fused)
            if (nsteps - t >= 2 && x2d_ok(g, fl, NC)) {
                if (run == 0) {
                    run_t = t;
                    for (int c = 0; c < NC; ++c) {
                        run_cur[c] = fl[c]->cur;
                        run_nstep[c] = fl[c]->nstep;
                    }
                }
                if (t + 2 == nsteps) last2 = x2d_replay(g, fl, forces);
                CHK(launch_x2d(g, fl, forces, ++run));
                for (int c = 0; c < NC; ++c) {
                    fl[c]->cur ^= 1;
                    advance_ueq(fl[c], 2);
                }
                t += 2;
                continue;
            }
            if (run) {
                unsigned bad = 0;
                CHK(x2d_failed(&bad));
                if (bad) {   // repeat from the input of launch `bad`, one step per launch
                    const int j = (int)bad - 1;
                    t = run_t + 2 * j;
                    for (int c = 0; c < NC; ++c) {
                        fl[c]->cur = run_cur[c] ^ (j & 1);
                        fl[c]->nstep = run_nstep[c] + 2u * (unsigned)j;
                        fl[c]->dilute = true;
                    }
                } else if (last2 && t == nsteps) {
                    for (int c = 0; c < NC; ++c) fl[c]->replay = last2;
                }
                last2.reset();
                run = 0;
                continue;
            }
            if (nsteps - t >= 3 && x2_ok(g, fl, NC) && x3_wanted(g, prec)) {
                const int rc = prec == SCL_FP32_DEV ? launch_x3<StoreF32>(g, fl, forces)
                                                    : launch_x3<StoreF16>(g, fl, forces);
                if (rc != SCL_OK) return rc;
                for (int c = 0; c < NC; ++c) {
                    fl[c]->cur ^= 1;
                    advance_ueq(fl[c], 3);
                }
                t += 3;
                continue;
            }
            if (nsteps - t >= 2 && x2_ok(g, fl, NC)) {
                const int rc = prec == SCL_FP32_DEV ? launch_x2<StoreF32>(g, fl, forces)
                                                    : launch_x2<StoreF16>(g, fl, forces);
                if (rc != SCL_OK) return rc;
                for (int c = 0; c < NC; ++c) {
                    fl[c]->cur ^= 1;
                    advance_ueq(fl[c], 2);
                }
                t += 2;
                continue;
            }
        }
        const bool gather = fl[0]->phase == PH_P;
        std::shared_ptr<Replay> rec;
        const bool last = t == nsteps - 1;
        int rc;
        switch (prec) {
            case SCL_FP64: rc = launch_fused<NC, StoreF64>(g, fl, forces, gather, last ? &rec : nullptr); break;
            case SCL_FP32_DEV: rc = launch_fused<NC, StoreF32>(g, fl, forces, gather, nullptr); break;
            default: rc = launch_fused<NC, StoreF16>(g, fl, forces, gather, nullptr); break;
        }
        if (rc != SCL_OK) return rc;
        for (int c = 0; c < NC; ++c) {
            fl[c]->cur ^= 1;
            fl[c]->phase = PH_P;
            set_pgeom(fl[c], const_cast<scl_geometry*>(g));
            if (rec) fl[c]->replay = rec;
            advance_ueq(fl[c], 1);
        }
        t += 1;
    }
    return SCL_OK;
}

}  // namespace

namespace scl {
int runtime_ready() { return lib_ready(); }
int runtime_device() { return g_lib.device; }
namespace serve {
int report(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
}  // namespace serve
}  // namespace scl

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char* scl_last_error(void) { return g_err.c_str(); }

int scl_init(int device) {
    if (g_lib.ready && g_lib.device != device)
        return fail(SCL_ERR_ARG, "library already initialised on device %d", g_lib.device);
    g_lib.device = device;
    return lib_ready();
}

int scl_synchronize(void) {
    CHK(lib_ready());
    HIPCHK(hipStreamSynchronize(g_lib.stream));
    return SCL_OK;
}

int scl_geometry_create(int nx, int ny, const int* mask, scl_geometry** out) {
    if (!out) return fail(SCL_ERR_ARG, "null output");
    *out = nullptr;
    if (nx <= 0 || ny <= 0 || !mask) return fail(SCL_ERR_ARG, "bad geometry size %dx%d or null mask", nx, ny);
    if ((long long)nx * ny * 9 > (1ll << 31) - 1) return fail(SCL_ERR_ARG, "lattice %dx%d too large", nx, ny);
    CHK(lib_ready());
    auto g = std::unique_ptr<scl_geometry>(new scl_geometry);
    g->nx = nx;
    g->ny = ny;
    g->pitch = (nx + 63) / 64 * 64;
    g->wpr = g->pitch / 64;
    g->plane = (long long)g->pitch * ny;
    g->host_mask.assign(mask, mask + (size_t)nx * ny);
    g->open = std::all_of(g->host_mask.begin(), g->host_mask.end(), [](int v) { return v == 1; });
    std::vector<uint64_t> bits((size_t)g->wpr * ny, 0ull);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
            if (mask[(size_t)y * nx + x] == 1) bits[(size_t)y * g->wpr + (x >> 6)] |= 1ull << (x & 63);
    // FNV-1a over the size and the packed words: a checkpoint's populations
    // (bounce-back slots included) belong to exactly this mask
    uint64_t hsh = 1469598103934665603ull;
    auto mix = [&hsh](uint64_t w) {
        for (int b = 0; b < 8; ++b) {
            hsh ^= (w >> (8 * b)) & 0xff;
            hsh *= 1099511628211ull;
        }
    };
    mix((uint64_t)nx);
    mix((uint64_t)ny);
    for (uint64_t w : bits) mix(w);
    g->mask_hash = hsh;
    CHK(dalloc(&g->mask, bits.size()));
    HIPCHK(hipMemcpy(g->mask, bits.data(), bits.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    CHK(dalloc(&g->part, (size_t)3 * ny + 3));
    *out = g.release();
    return SCL_OK;
}

static void geometry_free(scl_geometry* g) {
    if (g_lib.ready) (void)hipStreamSynchronize(g_lib.stream);
    dfree(g->mask);
    dfree(g->part);
    dfree(g->obs);
    delete g;
}

namespace {
void geometry_release(scl_geometry* g) {
    if (--g->refs == 0 && g->dead) geometry_free(g);
}
}  // namespace

int scl_geometry_destroy(scl_geometry* g) {
    if (!g) return SCL_OK;
    if (g->refs > 0) {
        g->dead = true;   // freed by the last field that uses it
        return SCL_OK;
    }
    geometry_free(g);
    return SCL_OK;
}

int scl_set_tile(scl_geometry* g, int cols, int rows) {
    if (!g) return fail(SCL_ERR_ARG, "null geometry");
    if (cols != 0 && cols != 64 && cols != 128 && cols != 256 && cols != 512)
        return fail(SCL_ERR_ARG, "tile columns must be 64, 128, 256 or 512 (got %d)", cols);
    if (rows < 0) return fail(SCL_ERR_ARG, "tile rows must be >= 0");
    g->tile_cols = cols;
    g->tile_rows = rows;
    return SCL_OK;
}

int scl_set_lookahead(scl_geometry* g, int rows_ahead) {
    if (!g) return fail(SCL_ERR_ARG, "null geometry");
    if (rows_ahead < 0 || rows_ahead > 2) return fail(SCL_ERR_ARG, "lookahead must be 0 (automatic), 1 or 2");
    g->lookahead = rows_ahead;
    return SCL_OK;
}

int scl_set_step_fusion(scl_geometry* g, int steps_per_launch, int rows) {
    if (!g) return fail(SCL_ERR_ARG, "null geometry");
    if (steps_per_launch < 0 || steps_per_launch > 3)
        return fail(SCL_ERR_ARG, "steps per launch must be 0 (automatic), 1, 2 or 3");
    if (rows < 0 || rows > 4096) return fail(SCL_ERR_ARG, "rows must be in 0..4096");
    g->fusion = steps_per_launch;
    g->x2_rows = rows;
    return SCL_OK;
}

int scl_set_fp64_fusion_watch(scl_geometry* g, double rho) {
    if (!g) return fail(SCL_ERR_ARG, "null geometry");
    if (!(rho >= kUeqWatch64) || !(rho <= 1e300)) return fail(SCL_ERR_ARG, "watch density must be finite and >= 1e-8");
    g->x2d_watch = rho;
    return SCL_OK;
}

int scl_get_step_fusion(const scl_geometry* g, int precision, int* steps_per_launch, int* rows) {
    if (!g || !steps_per_launch || !rows) return fail(SCL_ERR_ARG, "null argument");
    if (precision < SCL_FP64 || precision > SCL_FP16_DEV) return fail(SCL_ERR_ARG, "bad precision");
    if (precision == SCL_FP64) {
        const bool two = x2d_wanted(g);
        *steps_per_launch = two ? 2 : 1;
        *rows = two ? x2_rows_for(g, precision) : 0;
        return SCL_OK;
    }
    if (x3_wanted(g, precision)) {
        *steps_per_launch = 3;
        *rows = x3_rows_for(g, precision);
        return SCL_OK;
    }
    const bool two = x2_wanted(g, precision);
    *steps_per_launch = two ? 2 : 1;
    *rows = two ? x2_rows_for(g, precision) : 0;
    return SCL_OK;
}

int scl_get_tile(const scl_geometry* g, int ncomp, int precision, int* cols, int* rows, int* lookahead) {
    if (!g || !cols || !rows || !lookahead) return fail(SCL_ERR_ARG, "null argument");
    if (ncomp < 1 || ncomp > 2 || precision < SCL_FP64 || precision > SCL_FP16_DEV)
        return fail(SCL_ERR_ARG, "bad component count or precision");
    const Shape s = resolve_shape(g, ncomp, precision);
    *cols = 64 * s.nw;
    *rows = s.rows;
    *lookahead = s.la;
    return SCL_OK;
}

int scl_field_create(const scl_geometry* g, int precision, double tau, double rho0, scl_field** out) {
    if (!out) return fail(SCL_ERR_ARG, "null output");
    *out = nullptr;
    if (!g) return fail(SCL_ERR_ARG, "null geometry");
    if (precision < SCL_FP64 || precision > SCL_FP16_DEV) return fail(SCL_ERR_ARG, "bad precision %d", precision);
    CHK(lib_ready());
    auto f = std::unique_ptr<scl_field>(new scl_field);
    if (g->dead) return fail(SCL_ERR_ARG, "geometry already destroyed");
    f->g = const_cast<scl_geometry*>(g);
    f->precision = precision;
    f->tau = tau;
    f->omega = 1 / tau;              // include/fluid.h:62
    f->omc = 1.0 - f->omega;         // src/lbm.cpp:284, 308
    f->inv_omega = 1.0 / f->omega;   // src/lbm.cpp:213, 226
    // the kernels hold rho0 in FP32 and the conversions in FP64: one value
    // representable in both, so the encoding and the kernels agree
    f->rho0 = precision == SCL_FP64 ? 0.0 : (double)(float)rho0;
    const size_t bytes = n_pop(g) * f->esize();
    for (int b = 0; b < 2; ++b) {
        HIPCHK(hipMalloc(&f->f[b], bytes));
        HIPCHK(hipMemsetAsync(f->f[b], 0, bytes, g_lib.stream));
    }
    if (precision != SCL_FP64) {
        // zero populations in deviation form: d_i = -w_i*rho0 on every site
        CHK(ensure_stage(f.get()));
        HIPCHK(hipMemsetAsync(f->stage, 0, n_pop(g) * sizeof(double), g_lib.stream));
        const Grid2 gr = g->grid();
        if (precision == SCL_FP32_DEV)
            hipLaunchKernelGGL(k_from_f64<StoreF32>, site_grid(g), dim3(256), 0, g_lib.stream, gr, f->stage, f->rho0,
                               (float*)f->f[0]);
        else
            hipLaunchKernelGGL(k_from_f64<StoreF16>, site_grid(g), dim3(256), 0, g_lib.stream, gr, f->stage, f->rho0,
                               (_Float16*)f->f[0]);
        HIPCHK(hipGetLastError());
    }
    const size_t n = n_sites(g);
    if (precision == SCL_FP64) {
        CHK(dalloc(&f->ueqx, n));
        CHK(dalloc(&f->ueqy, n));
        HIPCHK(hipMemsetAsync(f->ueqx, 0, n * sizeof(double), g_lib.stream));
        HIPCHK(hipMemsetAsync(f->ueqy, 0, n * sizeof(double), g_lib.stream));
    } else {
        CHK(dalloc(&f->ueq_block, 2 * kUeqRing * n));
        HIPCHK(hipMemsetAsync(f->ueq_block, 0, 2 * kUeqRing * n * sizeof(double), g_lib.stream));
        for (int k = 0; k < kUeqRing; ++k)
            for (int d = 0; d < 2; ++d) f->ueqs[k][d] = f->ueq_block + (size_t)(2 * k + d) * n;
        f->ueqx = f->ueqs[0][0];
        f->ueqy = f->ueqs[0][1];
    }
    if (precision == SCL_FP64) {
        CHK(dalloc(&f->stamp, n + 2));
        HIPCHK(hipMemsetAsync(f->stamp, 0, (n + 2) * sizeof(unsigned), g_lib.stream));
    }
    f->g->refs++;
    *out = f.release();
    return SCL_OK;
}

int scl_field_destroy(scl_field* f) {
    if (!f) return SCL_OK;
    if (g_lib.ready) (void)hipStreamSynchronize(g_lib.stream);
    if (f->replay) {
        std::shared_ptr<Replay> r = f->replay;
        for (int c = 0; c < r->nc; ++c)
            if (r->fl[c] && r->fl[c] != f) {
                (void)materialize(r->fl[c]);
            }
        for (int c = 0; c < r->nc; ++c)
            if (r->fl[c]) r->fl[c]->replay.reset();
        (void)hipStreamSynchronize(g_lib.stream);
    }
    for (void* p : {f->f[0], f->f[1]}) dfree(p);
    if (f->precision == SCL_FP64) {
        dfree(f->ueqx);
        dfree(f->ueqy);
    } else {
        dfree(f->ueq_block);
    }
    for (double* p : {f->rho, f->ux, f->uy, f->frx, f->fry, f->fsx, f->fsy, f->feq, f->stage}) dfree(p);
    dfree(f->stamp);
    scl_geometry* g = f->g;
    scl_geometry* pg = f->pg;
    delete f;
    if (pg) geometry_release(pg);
    geometry_release(g);
    return SCL_OK;
}

int scl_field_set_interactions(scl_field* f, double g_solid, double g_fluid) {
    if (!f) return fail(SCL_ERR_ARG, "null field");
    f->g_solid = g_solid;
    f->g_fluid = g_fluid;
    return SCL_OK;
}

static double* aux_ptr(scl_field* f, int which) {
    switch (which) {
        case SCL_FEQ: return f->feq;
        case SCL_RHO: return f->rho;
        case SCL_UX: return f->ux;
        case SCL_UY: return f->uy;
        case SCL_UEQX: return f->ueqx;
        case SCL_UEQY: return f->ueqy;
        case SCL_FRX: return f->frx;
        case SCL_FRY: return f->fry;
        case SCL_FSX: return f->fsx;
        case SCL_FSY: return f->fsy;
        default: return nullptr;
    }
}

static int copy_planes(const scl_geometry* g, void* dev, const double* host, int planes, bool up) {
    const size_t spitch = (size_t)g->nx * sizeof(double), dpitch = (size_t)g->pitch * sizeof(double);
    for (int i = 0; i < planes; ++i) {
        double* d = (double*)dev + (size_t)i * g->plane;
        const double* h = host + (size_t)i * g->nx * g->ny;
        if (up)
            HIPCHK(hipMemcpy2DAsync(d, dpitch, h, spitch, spitch, g->ny, hipMemcpyHostToDevice, g_lib.stream));
        else
            HIPCHK(hipMemcpy2DAsync((void*)h, spitch, d, dpitch, spitch, g->ny, hipMemcpyDeviceToHost, g_lib.stream));
    }
    return SCL_OK;
}

int scl_field_upload(scl_field* f, int which, const double* src, size_t n) {
    if (!f || !src) return fail(SCL_ERR_ARG, "null argument");
    const scl_geometry* g = f->g;
    const size_t ns = (size_t)g->nx * g->ny;
    if (which == SCL_F) {
        if (n != 9 * ns) return fail(SCL_ERR_ARG, "populations need %zu values, got %zu", 9 * ns, n);
        CHK(materialize(f));
        if (f->precision == SCL_FP64) {
            CHK(copy_planes(g, f->f[f->cur], src, 9, true));
        } else {
            CHK(ensure_stage(f));
            CHK(copy_planes(g, f->stage, src, 9, true));
            const Grid2 gr = g->grid();
            if (f->precision == SCL_FP32_DEV)
                hipLaunchKernelGGL(k_from_f64<StoreF32>, site_grid(g), dim3(256), 0, g_lib.stream, gr, f->stage,
                                   f->rho0, (float*)f->f[f->cur]);
            else
                hipLaunchKernelGGL(k_from_f64<StoreF16>, site_grid(g), dim3(256), 0, g_lib.stream, gr, f->stage,
                                   f->rho0, (_Float16*)f->f[f->cur]);
            HIPCHK(hipGetLastError());
        }
        f->phase = PH_S;
        f->dilute = false;
        CHK(set_ueq_valid(f, f->nstep));
        HIPCHK(hipStreamSynchronize(g_lib.stream));
        return SCL_OK;
    }
    if (which < SCL_FEQ || which > SCL_FSY) return fail(SCL_ERR_ARG, "bad field selector %d", which);
    const size_t want = which == SCL_FEQ ? 9 * ns : ns;
    if (n != want) return fail(SCL_ERR_ARG, "field %d needs %zu values, got %zu", which, want, n);
    if (which != SCL_UEQX && which != SCL_UEQY) {
        if (f->precision != SCL_FP64) return fail(SCL_ERR_UNSUPPORTED, "field %d needs an SCL_FP64 field", which);
        CHK(ensure_aux(f));
    }
    CHK(materialize(f));
    CHK(copy_planes(g, aux_ptr(f, which), src, which == SCL_FEQ ? 9 : 1, true));
    if (which == SCL_UEQX || which == SCL_UEQY) CHK(set_ueq_valid(f, f->nstep));
    if (which == SCL_FSX || which == SCL_FSY) {
        // the reference's first F_solid entries are in use from now on
        f->has_solid_force = true;
        f->solid_force_uploaded = true;
    }
    HIPCHK(hipStreamSynchronize(g_lib.stream));
    return SCL_OK;
}

int scl_field_download(scl_field* f, int which, double* dst, size_t n) {
    if (!f || !dst) return fail(SCL_ERR_ARG, "null argument");
    const scl_geometry* g = f->g;
    const size_t ns = (size_t)g->nx * g->ny;
    if (which == SCL_F) {
        if (n != 9 * ns) return fail(SCL_ERR_ARG, "populations need %zu values, got %zu", 9 * ns, n);
        if (f->precision == SCL_FP64 && f->phase == PH_S) {
            CHK(copy_planes(g, f->f[f->cur], dst, 9, false));
        } else {
            CHK(ensure_stage(f));
            const bool gth = f->phase == PH_P;
            const Grid2 gr = (gth && f->pg ? f->pg : g)->grid();
            const void* s = f->f[f->cur];
            switch (f->precision) {
                case SCL_FP64:
                    hipLaunchKernelGGL((k_to_f64<StoreF64, true>), site_grid(g), dim3(256), 0, g_lib.stream, gr,
                                       (const double*)s, 0.0, f->stage);
                    break;
                case SCL_FP32_DEV:
                    if (gth)
                        hipLaunchKernelGGL((k_to_f64<StoreF32, true>), site_grid(g), dim3(256), 0, g_lib.stream, gr,
                                           (const float*)s, f->rho0, f->stage);
                    else
                        hipLaunchKernelGGL((k_to_f64<StoreF32, false>), site_grid(g), dim3(256), 0, g_lib.stream, gr,
                                           (const float*)s, f->rho0, f->stage);
                    break;
                default:
                    if (gth)
                        hipLaunchKernelGGL((k_to_f64<StoreF16, true>), site_grid(g), dim3(256), 0, g_lib.stream, gr,
                                           (const _Float16*)s, f->rho0, f->stage);
                    else
                        hipLaunchKernelGGL((k_to_f64<StoreF16, false>), site_grid(g), dim3(256), 0, g_lib.stream, gr,
                                           (const _Float16*)s, f->rho0, f->stage);
                    break;
            }
            HIPCHK(hipGetLastError());
            CHK(copy_planes(g, f->stage, dst, 9, false));
        }
        HIPCHK(hipStreamSynchronize(g_lib.stream));
        return check_ueq_flag(f);
    }
    if (which < SCL_FEQ || which > SCL_FSY) return fail(SCL_ERR_ARG, "bad field selector %d", which);
    const size_t want = which == SCL_FEQ ? 9 * ns : ns;
    if (n != want) return fail(SCL_ERR_ARG, "field %d needs %zu values, got %zu", which, want, n);
    if (which != SCL_UEQX && which != SCL_UEQY) {
        if (f->precision != SCL_FP64) return fail(SCL_ERR_UNSUPPORTED, "field %d needs an SCL_FP64 field", which);
        CHK(ensure_aux(f));
    }
    CHK(materialize(f));
    CHK(copy_planes(g, aux_ptr(f, which), dst, which == SCL_FEQ ? 9 : 1, false));
    HIPCHK(hipStreamSynchronize(g_lib.stream));
    return check_ueq_flag(f);
}

// ---- exact checkpoint -------------------------------------------------------

namespace {
// Checkpoint layout: this header, the populations as stored on the device
// (9 pitched planes in the storage form), then the equilibrium velocities
// the next step reads (x, then y; pitched FP64).
struct CkptHeader {
    char magic[8];
    int32_t precision, nx, ny, phase;
    double rho0;
    uint64_t pop_bytes, ueq_bytes;
    uint64_t mask_hash;                 // the geometry the populations belong to
};
constexpr char kCkptMagic[8] = {'S', 'C', 'L', 'C', 'K', 'P', 'T', '2'};

size_t ckpt_pop_bytes(const scl_field* f) { return n_pop(f->g) * f->esize(); }
size_t ckpt_ueq_bytes(const scl_field* f) { return 2 * n_sites(f->g) * sizeof(double); }
size_t ckpt_bytes(const scl_field* f) { return sizeof(CkptHeader) + ckpt_pop_bytes(f) + ckpt_ueq_bytes(f); }
}  // namespace

int scl_field_checkpoint_size(const scl_field* f, size_t* bytes) {
    if (!f || !bytes) return fail(SCL_ERR_ARG, "null argument");
    *bytes = ckpt_bytes(f);
    return SCL_OK;
}

int scl_field_checkpoint(scl_field* f, void* dst, size_t bytes) {
    if (!f || !dst) return fail(SCL_ERR_ARG, "null argument");
    if (bytes != ckpt_bytes(f)) return fail(SCL_ERR_ARG, "checkpoint needs %zu bytes, got %zu", ckpt_bytes(f), bytes);
    CHK(materialize(f));
    // post-collision populations of another mask: stream them with it first
    if (f->phase == PH_P && f->pg && f->pg != f->g) CHK(ensure_s(f));
    CkptHeader h{};
    std::memcpy(h.magic, kCkptMagic, sizeof h.magic);
    h.precision = f->precision;
    h.nx = f->g->nx;
    h.ny = f->g->ny;
    h.phase = f->phase;
    h.rho0 = f->rho0;
    h.mask_hash = f->g->mask_hash;
    h.pop_bytes = ckpt_pop_bytes(f);
    h.ueq_bytes = ckpt_ueq_bytes(f);
    char* out = (char*)dst;
    std::memcpy(out, &h, sizeof h);
    out += sizeof h;
    HIPCHK(hipMemcpyAsync(out, f->f[f->cur], h.pop_bytes, hipMemcpyDeviceToHost, g_lib.stream));
    out += h.pop_bytes;
    const size_t half = h.ueq_bytes / 2;
    HIPCHK(hipMemcpyAsync(out, f->ueqx, half, hipMemcpyDeviceToHost, g_lib.stream));
    HIPCHK(hipMemcpyAsync(out + half, f->ueqy, half, hipMemcpyDeviceToHost, g_lib.stream));
    HIPCHK(hipStreamSynchronize(g_lib.stream));
    return check_ueq_flag(f);
}

int scl_field_restore(scl_field* f, const void* src, size_t bytes) {
    if (!f || !src) return fail(SCL_ERR_ARG, "null argument");
    if (bytes != ckpt_bytes(f)) return fail(SCL_ERR_ARG, "checkpoint of %zu bytes for a field that needs %zu", bytes,
                                            ckpt_bytes(f));
    CkptHeader h;
    std::memcpy(&h, src, sizeof h);
    if (std::memcmp(h.magic, kCkptMagic, sizeof h.magic) != 0) return fail(SCL_ERR_ARG, "not a field checkpoint");
    if (h.precision != f->precision || h.nx != f->g->nx || h.ny != f->g->ny)
        return fail(SCL_ERR_ARG, "checkpoint of a %dx%d field of precision %d, field is %dx%d precision %d", h.nx, h.ny,
                    h.precision, f->g->nx, f->g->ny, f->precision);
    if (std::memcmp(&h.rho0, &f->rho0, sizeof(double)) != 0)
        return fail(SCL_ERR_ARG, "checkpoint deviations are about rho0 = %.17g, the field's about %.17g", h.rho0,
                    f->rho0);
    if (h.pop_bytes != ckpt_pop_bytes(f) || h.ueq_bytes != ckpt_ueq_bytes(f) || (h.phase != PH_S && h.phase != PH_P))
        return fail(SCL_ERR_ARG, "corrupt checkpoint header");
    if (h.mask_hash != f->g->mask_hash)
        return fail(SCL_ERR_ARG, "checkpoint of a field on another geometry (mask hash %016llx, field's %016llx)",
                    (unsigned long long)h.mask_hash, (unsigned long long)f->g->mask_hash);
    CHK(materialize(f));
    const char* in = (const char*)src + sizeof h;
    HIPCHK(hipMemcpyAsync(f->f[f->cur], in, h.pop_bytes, hipMemcpyHostToDevice, g_lib.stream));
    in += h.pop_bytes;
    const size_t half = h.ueq_bytes / 2;
    HIPCHK(hipMemcpyAsync(f->ueqx, in, half, hipMemcpyHostToDevice, g_lib.stream));
    HIPCHK(hipMemcpyAsync(f->ueqy, in + half, half, hipMemcpyHostToDevice, g_lib.stream));
    f->phase = h.phase;
    if (f->phase == PH_P) set_pgeom(f, f->g);
    CHK(set_ueq_valid(f, f->nstep));
    HIPCHK(hipStreamSynchronize(g_lib.stream));
    return SCL_OK;
}

// ---- per-operator entry points -------------------------------------------

int scl_compute_density(scl_field* f) {
    CHK(prepare_op(f));
    hipLaunchKernelGGL(k_density, site_grid(f->g), dim3(256), 0, g_lib.stream, f->g->grid(),
                       (const double*)f->f[f->cur], f->rho);
    HIPCHK(hipGetLastError());
    return SCL_OK;
}

int scl_compute_velocities(const scl_geometry* g, scl_field* f) {
    CHK(check_geom(g, f));
    CHK(prepare_op(f));
    hipLaunchKernelGGL(k_velocities, site_grid(g), dim3(256), 0, g_lib.stream, g->grid(), (const double*)f->f[f->cur],
                       f->rho, f->ux, f->uy);
    HIPCHK(hipGetLastError());
    return SCL_OK;
}

int scl_compute_f_equilibrium_geom(const scl_geometry* g, scl_field* f) {
    CHK(scl_compute_density(f));
    CHK(scl_compute_velocities(g, f));
    hipLaunchKernelGGL(k_feq, site_grid(g), dim3(256), 0, g_lib.stream, g->grid(), f->rho, f->ux, f->uy, f->feq);
    HIPCHK(hipGetLastError());
    return SCL_OK;
}

int scl_compute_f_equilibrium(scl_field* f) {
    CHK(prepare_op(f));
    hipLaunchKernelGGL(k_feq, site_grid(f->g), dim3(256), 0, g_lib.stream, f->g->grid(), f->rho, f->ueqx, f->ueqy,
                       f->feq);
    HIPCHK(hipGetLastError());
    return SCL_OK;
}

int scl_solid_surface_force(const scl_geometry* g, scl_field* a, scl_field* b) {
    CHK(check_geom(g, a));
    CHK(check_geom(g, b));
    for (scl_field* f : {a, b}) {
        if (f->has_solid_force) continue;   // the reference appends; the first N entries stay in use
        f->gsneg_captured = -1.0 * f->g_solid;
        f->has_solid_force = true;
        if (f->precision != SCL_FP64) continue;   // deviation forms recompute it from the mask
        CHK(ensure_aux(f));
        CHK(materialize(f));
        hipLaunchKernelGGL(k_solid_force, site_grid(g), dim3(256), 0, g_lib.stream, g->grid(), f->gsneg_captured,
                           f->fsx, f->fsy);
        HIPCHK(hipGetLastError());
    }
    return SCL_OK;
}

int scl_fluid_repulsion(const scl_geometry* g, scl_field* a, scl_field* b) {
    CHK(check_geom(g, a));
    CHK(check_geom(g, b));
    CHK(prepare_op(a));
    CHK(prepare_op(b));
    hipLaunchKernelGGL(k_repulsion, site_grid(g), dim3(256), 0, g_lib.stream, g->grid(), a->rho, b->rho,
                       -1.0 * a->g_fluid, -1.0 * b->g_fluid, a->frx, a->fry, b->frx, b->fry);
    HIPCHK(hipGetLastError());
    return SCL_OK;
}

int scl_equilibrium_velocities(const scl_geometry* g, scl_field* a, scl_field* b) {
    CHK(check_geom(g, a));
    CHK(check_geom(g, b));
    if (!a->has_solid_force || !b->has_solid_force)
        return fail(SCL_ERR_RANGE, "fluid-solid force not computed (compute_solid_surface_force first)");
    CHK(prepare_op(a));
    CHK(prepare_op(b));
    EqVelComp A{(const double*)a->f[a->cur], a->rho, a->omega, a->inv_omega, a->frx, a->fry, a->fsx, a->fsy,
                a->ux, a->uy, a->ueqx, a->ueqy};
    EqVelComp B{(const double*)b->f[b->cur], b->rho, b->omega, b->inv_omega, b->frx, b->fry, b->fsx, b->fsy,
                b->ux, b->uy, b->ueqx, b->ueqy};
    hipLaunchKernelGGL(k_eq_velocities, site_grid(g), dim3(256), 0, g_lib.stream, g->grid(), A, B);
    HIPCHK(hipGetLastError());
    // every site's value is this step's now (dry sites keep theirs, as the reference)
    CHK(set_ueq_valid(a, a->nstep + 1));
    return set_ueq_valid(b, b->nstep + 1);
}

int scl_collide1(const scl_geometry* g, scl_field* f) {
    CHK(check_geom(g, f));
    CHK(scl_compute_f_equilibrium_geom(g, f));
    hipLaunchKernelGGL(k_bgk, site_grid(g), dim3(256), 0, g_lib.stream, g->grid(), (double*)f->f[f->cur], f->feq,
                       f->omc, f->omega);
    HIPCHK(hipGetLastError());
    return SCL_OK;
}

int scl_collide2(scl_field* a, scl_field* b) {
    if (!a || !b) return fail(SCL_ERR_ARG, "null field");
    CHK(check_geom(a->g, b));
    CHK(scl_compute_f_equilibrium(a));
    CHK(scl_compute_f_equilibrium(b));
    for (scl_field* f : {a, b}) {
        hipLaunchKernelGGL(k_bgk, site_grid(f->g), dim3(256), 0, g_lib.stream, f->g->grid(), (double*)f->f[f->cur],
                           f->feq, f->omc, f->omega);
        HIPCHK(hipGetLastError());
    }
    return SCL_OK;
}

int scl_volume_force(const scl_geometry* g, scl_field* f, const double force[9]) {
    CHK(check_geom(g, f));
    if (!force) return fail(SCL_ERR_ARG, "null force");
    CHK(prepare_op(f));
    Force9 F;
    for (int i = 0; i < Q; ++i) F.v[i] = force[i];
    hipLaunchKernelGGL(k_volume_force, site_grid(g), dim3(256), 0, g_lib.stream, g->grid(), (double*)f->f[f->cur], F);
    HIPCHK(hipGetLastError());
    return SCL_OK;
}

static int stream_one(const scl_geometry* g, scl_field* f) {
    CHK(prepare_op(f));
    hipLaunchKernelGGL(k_gather<StoreF64>, site_grid(g), dim3(256), 0, g_lib.stream, g->grid(),
                       (const double*)f->f[f->cur], (double*)f->f[f->cur ^ 1]);
    HIPCHK(hipGetLastError());
    f->cur ^= 1;
    f->nstep++;
    return SCL_OK;
}

int scl_stream1(const scl_geometry* g, scl_field* f) {
    CHK(check_geom(g, f));
    return stream_one(g, f);
}

int scl_stream2(const scl_geometry* g, scl_field* a, scl_field* b) {
    CHK(check_geom(g, a));
    CHK(check_geom(g, b));
    CHK(stream_one(g, a));
    return stream_one(g, b);
}

// ---- fused steps ----------------------------------------------------------

int scl_step1(const scl_geometry* g, scl_field* f, const double* force, int nsteps) {
    if (!g || !f) return fail(SCL_ERR_ARG, "null argument");
    scl_field* fl[1] = {f};
    const double* forces[1] = {force};
    return fused_steps<1>(g, fl, forces, nsteps);
}

int scl_step2(const scl_geometry* g, scl_field* a, scl_field* b, const double* force_a, const double* force_b,
              int nsteps) {
    if (!g || !a || !b) return fail(SCL_ERR_ARG, "null argument");
    scl_field* fl[2] = {a, b};
    const double* forces[2] = {force_a, force_b};
    return fused_steps<2>(g, fl, forces, nsteps);
}

// ---- diagnostics ----------------------------------------------------------

namespace {
// FP64 populations of f's streamed state in f->stage (gathered with the mask
// that produced them if f holds post-collision populations); *gr: that grid.
int stage_f64(const scl_geometry* g, scl_field* f, Grid2* gr_out) {
    CHK(ensure_stage(f));
    const void* s = f->f[f->cur];
    const bool gth = f->phase == PH_P;
    const Grid2 gr = (gth && f->pg ? f->pg : g)->grid();
    switch (f->precision) {
        case SCL_FP64:
            if (gth) hipLaunchKernelGGL((k_to_f64<StoreF64, true>), site_grid(g), dim3(256), 0, g_lib.stream, gr, (const double*)s, 0.0, f->stage);
            else hipLaunchKernelGGL((k_to_f64<StoreF64, false>), site_grid(g), dim3(256), 0, g_lib.stream, gr, (const double*)s, 0.0, f->stage);
            break;
        case SCL_FP32_DEV:
            if (gth) hipLaunchKernelGGL((k_to_f64<StoreF32, true>), site_grid(g), dim3(256), 0, g_lib.stream, gr, (const float*)s, f->rho0, f->stage);
            else hipLaunchKernelGGL((k_to_f64<StoreF32, false>), site_grid(g), dim3(256), 0, g_lib.stream, gr, (const float*)s, f->rho0, f->stage);
            break;
        default:
            if (gth) hipLaunchKernelGGL((k_to_f64<StoreF16, true>), site_grid(g), dim3(256), 0, g_lib.stream, gr, (const _Float16*)s, f->rho0, f->stage);
            else hipLaunchKernelGGL((k_to_f64<StoreF16, false>), site_grid(g), dim3(256), 0, g_lib.stream, gr, (const _Float16*)s, f->rho0, f->stage);
            break;
    }
    HIPCHK(hipGetLastError());
    if (gr_out) *gr_out = gr;
    return SCL_OK;
}

// scl_moments without allocating: the populations are read as stored (and
// gathered on the fly) instead of through the field's FP64 staging buffer,
// and the partial sums go to the static `scratch` array instead of the
// geometry's.  Same per-site values and summation order as scl_moments, so
// the same bits.
int scl_moments_noalloc(const scl_geometry* g, scl_field* f, double out[3]) {
    CHK(check_geom(g, f));
    if (!out) return fail(SCL_ERR_ARG, "null output");
    if (g->ny > kScratchRows)
        return fail(SCL_ERR_ARG, "lattice of %d rows: the allocation-free moments hold partial sums for at most %d",
                    g->ny, kScratchRows);
    CHK(lib_ready());
    double* part = g_lib.scratch;
    const bool gth = f->phase == PH_P;
    const Grid2 gr = (gth && f->pg ? f->pg : g)->grid();
    const void* s = f->f[f->cur];
    const dim3 grid(g->ny), block(256);
    switch (f->precision) {
        case SCL_FP64:
            if (gth) hipLaunchKernelGGL((k_moments_direct<StoreF64, true>), grid, block, 0, g_lib.stream, gr, (const double*)s, 0.0, part);
            else hipLaunchKernelGGL((k_moments_direct<StoreF64, false>), grid, block, 0, g_lib.stream, gr, (const double*)s, 0.0, part);
            break;
        case SCL_FP32_DEV:
            if (gth) hipLaunchKernelGGL((k_moments_direct<StoreF32, true>), grid, block, 0, g_lib.stream, gr, (const float*)s, f->rho0, part);
            else hipLaunchKernelGGL((k_moments_direct<StoreF32, false>), grid, block, 0, g_lib.stream, gr, (const float*)s, f->rho0, part);
            break;
        default:
            if (gth) hipLaunchKernelGGL((k_moments_direct<StoreF16, true>), grid, block, 0, g_lib.stream, gr, (const _Float16*)s, f->rho0, part);
            else hipLaunchKernelGGL((k_moments_direct<StoreF16, false>), grid, block, 0, g_lib.stream, gr, (const _Float16*)s, f->rho0, part);
            break;
    }
    hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(256), 0, g_lib.stream, g->ny, (const double*)part,
                       part + 3 * (size_t)g->ny);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, part + 3 * (size_t)g->ny, 3 * sizeof(double), hipMemcpyDeviceToHost, g_lib.stream));
    HIPCHK(hipStreamSynchronize(g_lib.stream));
    return check_ueq_flag(f);
}

// Row extents of the two droplet-observable site sets (k_obs_rows) and the
// densities at sites s0, s1, on the host.
int droplet_rows(const scl_geometry* g, scl_field* d, scl_field* b, double rho_bulk, double rho_disp, int s0,
                 int s1, std::vector<int>* rows, double pts[4]) {
    if (!d || !b) return fail(SCL_ERR_ARG, "null field");
    if (d == b) return fail(SCL_ERR_ARG, "the droplet and bulk components must be different fields");
    CHK(check_geom(g, d));
    CHK(check_geom(g, b));
    Grid2 gr{};
    CHK(stage_f64(g, d, &gr));
    CHK(stage_f64(g, b, nullptr));
    if (!g->obs) CHK(dalloc(&const_cast<scl_geometry*>(g)->obs, (size_t)6 * g->ny));
    hipLaunchKernelGGL(k_obs_rows, dim3(g->ny), dim3(256), 0, g_lib.stream, gr, (const double*)d->stage,
                       (const double*)b->stage, rho_bulk, rho_disp, g->obs);
    hipLaunchKernelGGL(k_obs_points, dim3(1), dim3(1), 0, g_lib.stream, gr, (const double*)d->stage,
                       (const double*)b->stage, s0, s1, g->part);
    HIPCHK(hipGetLastError());
    rows->assign((size_t)6 * g->ny, 0);
    HIPCHK(hipMemcpyAsync(rows->data(), g->obs, rows->size() * sizeof(int), hipMemcpyDeviceToHost, g_lib.stream));
    HIPCHK(hipMemcpyAsync(pts, g->part, 4 * sizeof(double), hipMemcpyDeviceToHost, g_lib.stream));
    HIPCHK(hipStreamSynchronize(g_lib.stream));
    return SCL_OK;
}
}  // namespace

int scl_moments(const scl_geometry* g, scl_field* f, double out[3]) {
    CHK(check_geom(g, f));
    if (!out) return fail(SCL_ERR_ARG, "null output");
    Grid2 gr{};
    CHK(stage_f64(g, f, &gr));
    hipLaunchKernelGGL(k_moments_rows, dim3(g->ny), dim3(256), 0, g_lib.stream, gr, (const double*)f->stage, g->part);
    hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(256), 0, g_lib.stream, g->ny, (const double*)g->part,
                       g->part + 3 * (size_t)g->ny);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, g->part + 3 * (size_t)g->ny, 3 * sizeof(double), hipMemcpyDeviceToHost, g_lib.stream));
    HIPCHK(hipStreamSynchronize(g_lib.stream));
    return check_ueq_flag(f);
}

// Laplace law, as laplace_law/compute_surface_tension.py:63-70 evaluates it
// from the two density fields: the pressure p = (rho_b + rho_d)/3 +
// G/3 rho_b rho_d at two sites, and the radius of the droplet set.
int scl_laplace_observables(const scl_geometry* g, scl_field* droplet, scl_field* bulk, double G, int x_in,
                            int y_in, int x_out, int y_out, double rho_bulk, double rho_disp, double out[2]) {
    if (!g || !out) return fail(SCL_ERR_ARG, "null argument");
    if (x_in < 0 || x_in >= g->nx || y_in < 0 || y_in >= g->ny || x_out < 0 || x_out >= g->nx || y_out < 0 ||
        y_out >= g->ny)
        return fail(SCL_ERR_RANGE, "pressure sites outside the %dx%d lattice", g->nx, g->ny);
    std::vector<int> rows;
    double pts[4];
    CHK(droplet_rows(g, droplet, bulk, rho_bulk, rho_disp, y_in * g->pitch + x_in, y_out * g->pitch + x_out, &rows,
                     pts));
    int ylo = -1, yhi = -1, xlo = g->nx, xhi = -1;
    for (int y = 0; y < g->ny; ++y) {
        if (!rows[6 * y]) continue;
        if (ylo < 0) ylo = y;
        yhi = y;
        xlo = std::min(xlo, rows[6 * y + 1]);
        xhi = std::max(xhi, rows[6 * y + 2]);
    }
    if (ylo < 0) return fail(SCL_ERR_RANGE, "no site of the droplet set (rho_d < rho_b, rho_disp < rho_b <= rho_bulk)");
    auto p = [&](double rd, double rb) { return 1.0 / 3.0 * (rb + rd) + G / 3.0 * rb * rd; };
    out[0] = p(pts[0], pts[1]) - p(pts[2], pts[3]);
    // the script's [0] index is the row (y), [1] the column (x)
    const double ry = (yhi - ylo) / 2.0, rx = (xhi - xlo) / 2.0;
    out[1] = 0.5 * (ry + rx);
    return SCL_OK;
}

// Contact angle of a droplet on a solid below it, as
// contact_angles/compute_contact_angles.py:46-100 evaluates it: the base
// from the interface's two lowest rows (extrapolated half a row down), the
// height from its extent in y, the circle through them, and the angle.
int scl_contact_angle(const scl_geometry* g, scl_field* droplet, scl_field* bulk, double rho_bulk, double rho_disp,
                      double out[4]) {
    if (!g || !out) return fail(SCL_ERR_ARG, "null argument");
    std::vector<int> rows;
    double pts[4];
    CHK(droplet_rows(g, droplet, bulk, rho_bulk, rho_disp, 0, 0, &rows, pts));
    int ylo = -1, yhi = -1;
    for (int y = 0; y < g->ny; ++y)
        if (rows[6 * y + 3]) {
            if (ylo < 0) ylo = y;
            yhi = y;
        }
    if (ylo < 0) return fail(SCL_ERR_RANGE, "no interface site (rho_b <= rho_d, rho_d <= rho_bulk, rho_b > rho_disp)");
    if (ylo + 1 >= g->ny || !rows[6 * (ylo + 1) + 3])
        return fail(SCL_ERR_RANGE, "the interface has no site in the row above its lowest one");
    const double y2 = ylo, y1 = ylo + 1;                  // yL2 = yR2, yL1 = yR1
    double xl1 = rows[6 * (ylo + 1) + 4], xl2 = rows[6 * ylo + 4];
    double xr1 = rows[6 * (ylo + 1) + 5], xr2 = rows[6 * ylo + 5];
    const double yb = ylo - 0.5;                          // yL = yR
    if (xl1 == xl2) xl1 = xl1 - 1;
    if (xr1 == xr2) xr1 = xr1 - 1;
    const double sl = (y1 - y2) / (xl1 - xl2), sr = (y1 - y2) / (xr1 - xr2);
    const double il = y2 - sl * xl2, ir = y2 - sr * xr2;
    const double xl = 1 / sl * (yb - il), xr = 1 / sr * (yb - ir);
    const double L = xr - xl;
    const double H = (yhi - ylo) + 0.5;
    const double R = (4 * H * H + L * L) / (8 * H);
    double theta = std::atan(L / (2 * (R - H))) * (180.0 / M_PI);
    if (theta < 0) theta += 180;
    out[0] = theta;
    out[1] = L;
    out[2] = H;
    out[3] = R;
    return SCL_OK;
}

int scl_timing(int reset, double* avg_ms, int* count) {
    CHK(lib_ready());
    if (reset) {
        HIPCHK(hipStreamSynchronize(g_lib.stream));
        for (auto& p : g_lib.ev) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
        g_lib.ev.clear();
        g_lib.ev_steps.clear();
        g_lib.timing = reset > 0;
        if (avg_ms) *avg_ms = 0.0;
        if (count) *count = 0;
        return SCL_OK;
    }
    HIPCHK(hipStreamSynchronize(g_lib.stream));
    double tot = 0.0;
    long long steps = 0;
    for (size_t k = 0; k < g_lib.ev.size(); ++k) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, g_lib.ev[k].first, g_lib.ev[k].second));
        tot += ms;
        steps += g_lib.ev_steps[k];
    }
    if (avg_ms) *avg_ms = steps ? tot / steps : 0.0;
    if (count) *count = (int)g_lib.ev.size();
    return SCL_OK;
}

int scl_autotune(scl_geometry* g, int ncomp, int precision, int* cols, int* rows, int* lookahead) {
    if (!g) return fail(SCL_ERR_ARG, "null geometry");
    if (ncomp < 1 || ncomp > 2 || precision < SCL_FP64 || precision > SCL_FP16_DEV)
        return fail(SCL_ERR_ARG, "bad component count or precision");
    CHK(lib_ready());
    static const size_t esz[3] = {8, 4, 2};
    const bool wide = needs_wide(g, esz[precision]);
    // scratch fields at rest (their contents do not affect the timing)
    scl_field* fl[2] = {nullptr, nullptr};
    int rc = SCL_OK;
    for (int c = 0; c < ncomp && rc == SCL_OK; ++c) {
        rc = scl_field_create(g, precision, 1.0, 1.0, &fl[c]);
        if (rc == SCL_OK) rc = scl_field_set_interactions(fl[c], 0.0, 0.9);
        if (rc == SCL_OK) {
            const Grid2 gr = g->grid();
            void* d = fl[c]->f[fl[c]->cur];
            if (precision == SCL_FP64)
                hipLaunchKernelGGL(k_fill_rest<StoreF64>, site_grid(g), dim3(256), 0, g_lib.stream, gr, (double*)d, 1.0, 1.0);
            else if (precision == SCL_FP32_DEV)
                hipLaunchKernelGGL(k_fill_rest<StoreF32>, site_grid(g), dim3(256), 0, g_lib.stream, gr, (float*)d, 1.0, 1.0);
            else
                hipLaunchKernelGGL(k_fill_rest<StoreF16>, site_grid(g), dim3(256), 0, g_lib.stream, gr, (_Float16*)d, 1.0, 1.0);
            if (hipGetLastError() != hipSuccess) rc = fail(SCL_ERR_HIP, "fill kernel launch failed");
        }
    }
    if (rc == SCL_OK && ncomp == 2) rc = scl_solid_surface_force(g, fl[0], fl[1]);
    // candidates (the caller's own overrides are set aside while timing)
    const int saved[3] = {g->tile_cols, g->tile_rows, g->lookahead};
    const int saved_fusion = g->fusion, saved_x2 = g->x2_rows;
    g->tile_cols = g->tile_rows = g->lookahead = 0;
    g->fusion = 1;   // single-step shapes first
    int* slot = g->tuned[ncomp - 1][precision];
    const int saved_tuned[3] = {slot[0], slot[1], slot[2]};
    double best = 1e300;
    int bc = 0, br = 0, bl = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == SCL_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(SCL_ERR_HIP, "event creation failed");
    const int colset[4] = {64, 128, 256, 512};
    const int rowset[6] = {4, 8, 16, 32, 64, 128};
    const int pitch = g->pitch;
    for (int ci = 0; ci < 4 && rc == SCL_OK; ++ci) {
        const int c = colset[ci];
        if (wide && c != 256) continue;
        if (c > 2 * pitch && c > 64) continue;
        for (int ri = 0; ri < 6 && rc == SCL_OK; ++ri) {
            const int r = rowset[ri];
            if (r > g->ny && ri > 0) continue;
            for (int la = 1; la <= 2 && rc == SCL_OK; ++la) {
                if (wide && la != 1) continue;
                slot[0] = c;
                slot[1] = r;
                slot[2] = la;
                const double* none[2] = {nullptr, nullptr};
                rc = ncomp == 1 ? fused_steps<1>(g, fl, none, 2) : fused_steps<2>(g, fl, none, 2);
                if (rc != SCL_OK) break;
                int reps = 3;
                float ms = 0.f;
                for (int pass = 0; pass < 2 && rc == SCL_OK; ++pass) {
                    if (hipEventRecord(e0, g_lib.stream) != hipSuccess) rc = fail(SCL_ERR_HIP, "event record failed");
                    if (rc == SCL_OK) rc = ncomp == 1 ? fused_steps<1>(g, fl, none, reps) : fused_steps<2>(g, fl, none, reps);
                    if (rc == SCL_OK && (hipEventRecord(e1, g_lib.stream) != hipSuccess ||
                                         hipEventSynchronize(e1) != hipSuccess ||
                                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                        rc = fail(SCL_ERR_HIP, "event timing failed");
                    if (pass == 0 && rc == SCL_OK) {
                        const double per = ms / reps;
                        int want = per > 0 ? (int)(3.0 / per) : 50;   // about 3 ms per candidate
                        reps = want < 3 ? 3 : (want > 50 ? 50 : want);
                    }
                }
                if (rc == SCL_OK && ms / reps < best) {
                    best = ms / reps;
                    bc = c;
                    br = r;
                    bl = la;
                }
            }
        }
    }
    // two steps per launch: rows per band, kept only if faster than single
    // steps.  (FP64 calls end in a single step: odd step counts, so that
    // all but one go two per launch.)
    const bool f64 = precision == SCL_FP64;
    double best2 = 1e300;
    if (rc == SCL_OK && bc && ncomp == 2 && (f64 ? x2d_supported(g) : x2_supported(g, precision))) {
        slot[0] = bc;
        slot[1] = br;
        slot[2] = bl;
        g->fusion = 2;
        int b2 = 0;
        // whole waves of workgroups (1..4), and two fixed heights
        int bandset[6] = {16, 32, x2_wave_rows(g, precision, 1), x2_wave_rows(g, precision, 2),
                          x2_wave_rows(g, precision, 3), x2_wave_rows(g, precision, 4)};
        for (int bi = 0; bi < 6 && rc == SCL_OK; ++bi) {
            bool seen = bandset[bi] > g->ny && bi > 0;
            for (int bj = 0; bj < bi; ++bj) seen = seen || bandset[bj] == bandset[bi];
            if (seen) continue;
            g->x2_rows = bandset[bi];
            const double* none[2] = {nullptr, nullptr};
            rc = fused_steps<2>(g, fl, none, f64 ? 3 : 2);
            int reps = f64 ? 5 : 4;
            float ms = 0.f;
            for (int pass = 0; pass < 2 && rc == SCL_OK; ++pass) {
                if (hipEventRecord(e0, g_lib.stream) != hipSuccess) rc = fail(SCL_ERR_HIP, "event record failed");
                if (rc == SCL_OK) rc = fused_steps<2>(g, fl, none, reps);
                if (rc == SCL_OK && (hipEventRecord(e1, g_lib.stream) != hipSuccess ||
                                     hipEventSynchronize(e1) != hipSuccess ||
                                     hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                    rc = fail(SCL_ERR_HIP, "event timing failed");
                if (pass == 0 && rc == SCL_OK) {
                    const double per = ms / reps;
                    int want = per > 0 ? (int)(3.0 / per) : 50;
                    want = want < 4 ? 4 : (want > 50 ? 50 : want);
                    reps = f64 ? (want | 1) : (want & ~1);
                }
            }
            if (rc == SCL_OK && ms / reps < best2) {
                best2 = ms / reps;
                b2 = bandset[bi];
            }
        }
        if (rc == SCL_OK) g->x2_tuned[precision] = best2 < best ? b2 : -1;
    }
    // three steps per launch (deviation forms), kept only if faster than both
    if (rc == SCL_OK && bc && ncomp == 2 && !f64 && x2_supported(g, precision)) {
        g->fusion = 3;
        double best3 = 1e300;
        int b3 = 0;
        int bandset3[6] = {32, 64, x3_wave_rows(g, precision, 1), x3_wave_rows(g, precision, 2),
                           x3_wave_rows(g, precision, 3), x3_wave_rows(g, precision, 4)};
        for (int bi = 0; bi < 6 && rc == SCL_OK; ++bi) {
            bool seen = bandset3[bi] > g->ny && bi > 0;
            for (int bj = 0; bj < bi; ++bj) seen = seen || bandset3[bj] == bandset3[bi];
            if (seen) continue;
            g->x2_rows = bandset3[bi];
            const double* none[2] = {nullptr, nullptr};
            rc = fused_steps<2>(g, fl, none, 3);
            int reps = 6;
            float ms = 0.f;
            for (int pass = 0; pass < 2 && rc == SCL_OK; ++pass) {
                if (hipEventRecord(e0, g_lib.stream) != hipSuccess) rc = fail(SCL_ERR_HIP, "event record failed");
                if (rc == SCL_OK) rc = fused_steps<2>(g, fl, none, reps);
                if (rc == SCL_OK && (hipEventRecord(e1, g_lib.stream) != hipSuccess ||
                                     hipEventSynchronize(e1) != hipSuccess ||
                                     hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                    rc = fail(SCL_ERR_HIP, "event timing failed");
                if (pass == 0 && rc == SCL_OK) {
                    const double per = ms / reps;
                    int want = per > 0 ? (int)(3.0 / per) : 48;
                    want = want < 6 ? 6 : (want > 48 ? 48 : want);
                    reps = want / 3 * 3;
                }
            }
            if (rc == SCL_OK && ms / reps < best3) {
                best3 = ms / reps;
                b3 = bandset3[bi];
            }
        }
        if (rc == SCL_OK) g->x3_tuned[precision] = best3 < std::min(best, best2) ? b3 : -1;
    }
    g->fusion = saved_fusion;
    g->x2_rows = saved_x2;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    g->tile_cols = saved[0];
    g->tile_rows = saved[1];
    g->lookahead = saved[2];
    for (int c = 0; c < ncomp; ++c)
        if (fl[c]) (void)scl_field_destroy(fl[c]);
    if (rc != SCL_OK || !bc) {
        slot[0] = saved_tuned[0];
        slot[1] = saved_tuned[1];
        slot[2] = saved_tuned[2];
        return rc != SCL_OK ? rc : fail(SCL_ERR_ARG, "no launch shape could be timed");
    }
    slot[0] = bc;
    slot[1] = br;
    slot[2] = bl;
    if (cols) *cols = bc;
    if (rows) *rows = br;
    if (lookahead) *lookahead = bl;
    return SCL_OK;
}

int scl_probe_copy(size_t bytes, int reps, double* avg_ms) {
    if (!avg_ms || bytes < 16 || reps <= 0) return fail(SCL_ERR_ARG, "bad probe arguments");
    CHK(lib_ready());
    const size_t n = bytes / 16;
    f32x4 *a = nullptr, *b = nullptr;
    CHK(dalloc(&a, n));
    int rc = dalloc(&b, n);
    if (rc != SCL_OK) {
        dfree(a);
        return rc;
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemsetAsync(a, 0, n * 16, g_lib.stream));
    const unsigned nblk = (unsigned)((n + 1023) / 1024);
    hipLaunchKernelGGL(k_probe_copy, dim3(nblk), dim3(256), 0, g_lib.stream, a, b, n);
    HIPCHK(hipEventRecord(e0, g_lib.stream));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_probe_copy, dim3(nblk), dim3(256), 0, g_lib.stream, a, b, n);
    HIPCHK(hipEventRecord(e1, g_lib.stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(a);
    dfree(b);
    return SCL_OK;
}

}  // extern "C"
