// step.h -- fused D2Q9 Shan-Chen step for gfx950 (one or two components).
//
// One launch advances every site by one reference step.  It stores the
// post-collision populations P_n (pull scheme): the streamed state of the
// reference, S_{n+1}, is Gather(P_n), with Gather the link bounce-back pull
//   S_i(x) = P_i(x - c_i)          if x - c_i is fluid
//          = P_opp(i)(x)           otherwise,
// which is exactly the reference's push stream (src/lbm.cpp:343-435) seen
// from the receiving site.  The next launch performs that gather while it
// reads, so populations cross HBM once per step (read 9*NC, write 9*NC).
//
// Work decomposition: a workgroup owns a strip of BW = 64*NW columns and
// `rows` rows and walks it from south to north.  Each thread owns one
// column.  The densities of both components live in a 4-row LDS ring with
// one halo column at each side of the strip; the halo densities are
// recomputed from gathered populations, so no density array crosses HBM.
// Per row y the loop is software-pipelined: the loads of row y+2 are issued
// first (their addresses come from mask rows fetched one iteration earlier),
// row y is finished while they are in flight (forces from the three LDS rows
// y-1..y+1, the rest from registers), and the density of row y+2 enters the
// ring afterwards.  One barrier per row.  Loads and stores use 32-bit byte
// offsets from the component's base address, shared by both components.
//
// FP64 arithmetic reproduces the reference operation by operation (the file
// is compiled with -ffp-contract=off): density src/fluid.cpp:177-182,
// repulsion src/lbm.cpp:163-204, adhesion src/lbm.cpp:110-140 with
// src/fluid.cpp:154-160, equilibrium velocity src/lbm.cpp:250-269, feq
// src/fluid.cpp:254-276, BGK src/lbm.cpp:308-309, volume force
// src/lbm.cpp:335-336.  The deviation forms (FP32 / FP16 storage) compute
// the same physics on d_i = f_i - w_i*rho0 in FP32.
#pragma once
#include "lattice.h"

namespace scl {

template <class T>
struct CompArgs {
    const void* src;   // populations read this step (P_{n-1}, or S_n for the first step)
    void* dst;         // P_n
    T omega, omc, inv_omega;   // omega = 1/tau, omc = 1 - omega, inv_omega = 1/omega
    T gneg;            // -1.0 * G_fluid
    T gsneg;           // -1.0 * G_solid (as captured by the solid-force call)
    T rho0;            // deviation reference density (deviation forms)
    T force[9];
    int has_force;
    // FP64: index of the time step this launch computes, and the zero-density
    // rule's records (ueq_hold / ueq_reuse): stamp[s] = the step whose
    // equilibrium velocity the held value of site s is; stamp[plane] = error
    // flag; stamp[plane + 1] = the step as of which every site's held value is
    // valid (set by the host).  Unused by the deviation forms.
    unsigned step;
    unsigned* stamp;
    double* ueqx;      // equilibrium velocity store (dry-site fallback / replay output)
    double* ueqy;
    // Deviation forms: the equilibrium velocities of the previous step, read
    // at zero-density sites (rotating buffers; FP64 reads and writes ueqx/ueqy)
    const double* ueqx_in;
    const double* ueqy_in;
    // replay outputs (FP64, pitched), written only by the replay variant
    double* rho;
    double* jx;
    double* jy;
    double* frx;
    double* fry;
    double* feq;       // 9 planes
};

// The body force as the kernels take it (host side): F_i for FP64; for the
// deviation forms (relax_dev) F_i / omega for i > 0 and the sum of all F_i,
// the mass it adds, in slot 0.
template <class T>
inline void set_force(CompArgs<T>& c, const double* force, double omega) {
    c.has_force = force ? 1 : 0;
    if (!force) {
        for (int i = 0; i < Q; ++i) c.force[i] = T(0);
    } else if (sizeof(T) == sizeof(double)) {
        for (int i = 0; i < Q; ++i) c.force[i] = (T)force[i];
    } else {
        double sum = 0.0;
        for (int i = 0; i < Q; ++i) sum += force[i];
        c.force[0] = (T)sum;
        for (int i = 1; i < Q; ++i) c.force[i] = (T)(force[i] / omega);
    }
}

template <class T>
struct StepArgs {
    int nx, ny, pitch, wpr;
    long long plane;
    const uint64_t* mask;
    int rows;          // rows per workgroup
    int gx;            // workgroups along x
    CompArgs<T> c[2];
};

// Byte offsets of the 9 populations that site (col, r) reads: the pull
// sources, or its own opposite slot where the source is solid.  The choice
// is made with a bit mask rather than a branch, so the nine loads of a site
// issue back to back without divergent control flow between them.  GHOST:
// the source buffer holds the bounce-back values in the solid sources' slots
// (see ghost_stores), so every pull reads its source slot.
template <class ST, bool GATHER, class OffT, bool GHOST = false>
__device__ inline void pull_offsets(OffT plane_b, OffT pitch_b, int col, int colm, int colp, int r, int rm,
                                    int rp, unsigned nb, OffT (&off)[Q]) {
    constexpr OffT es = sizeof(typename ST::S);
    const OffT orr = (OffT)r * pitch_b, os = (OffT)rm * pitch_b, on = (OffT)rp * pitch_b;
    const OffT cc = (OffT)col * es, cm = (OffT)colm * es, cp = (OffT)colp * es;
    const OffT home = orr + cc;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        if (!GATHER || i == 0) {
            off[i] = (OffT)i * plane_b + home;
        } else {
            const OffT row = cy(i) == 1 ? os : (cy(i) == -1 ? on : orr);
            const OffT col_b = cx(i) == 1 ? cm : (cx(i) == -1 ? cp : cc);
            const OffT pull = (OffT)i * plane_b + row + col_b;
            if constexpr (GHOST) {
                off[i] = pull;
            } else {
                const OffT back = (OffT)opp(i) * plane_b + home;
                const OffT keep = (OffT)0 - (OffT)((nb >> opp(i)) & 1u);   // all ones where the source is fluid
                off[i] = back ^ ((pull ^ back) & keep);
            }
        }
    }
}

// Element-indexed form of the gather (used by the single-pass helper
// kernels): populations of site (x, r) with 64-bit offsets.
template <class ST, bool GATHER>
__device__ inline void load_site(const typename ST::S* __restrict__ src, long long plane, int pitch, int x,
                                 int xm, int xp, int r, int rm, int rp, unsigned nb, typename ST::T (&p)[Q]) {
    constexpr unsigned long long es = sizeof(typename ST::S);
    unsigned long long off[Q];
    pull_offsets<ST, GATHER, unsigned long long>((unsigned long long)plane * es, (unsigned long long)pitch * es,
                                                 x, xm, xp, r, rm, rp, nb, off);
#pragma unroll
    for (int i = 0; i < Q; ++i) p[i] = ST::ld((const typename ST::S*)((const char*)src + off[i]));
}

template <class ST, class OffT>
__device__ inline typename ST::T ld_at(const void* base, OffT off) {
    return ST::ld((const typename ST::S*)((const char*)base + off));
}
// in working units (ST::kUnit)
template <class ST, class OffT>
__device__ inline typename ST::T ldw_at(const void* base, OffT off) {
    return ST::ldw((const typename ST::S*)((const char*)base + off));
}

template <class ST, class OffT>
__device__ inline void st_at(void* base, OffT off, typename ST::T v) {
    ST::st((typename ST::S*)((char*)base + off), v);
}

// Density in the reference order, ((0.0 + f0) + f1) + ... (deviation forms:
// sum of deviations).
template <class T>
__device__ inline T dens(const T (&p)[Q]) {
    T r = T(0);
#pragma unroll
    for (int i = 0; i < Q; ++i) r += p[i];
    return r;
}
// FP32 (deviation forms): the same sum without the leading 0 + (which only
// turns a -0 sum into +0 and costs an instruction per site).
__device__ inline float dens(const float (&p)[Q]) {
    float r = p[0];
#pragma unroll
    for (int i = 1; i < Q; ++i) r += p[i];
    return r;
}
// The density (deviation) of populations in working units, in plain units:
// the sum in working units times 1/kUnit, exactly the sum of the decoded
// values (a power-of-two scale commutes with every rounding of the sum).
template <class ST>
__device__ inline typename ST::T dens_w(const typename ST::T (&p)[Q]) {
    if constexpr (ST::kUnit == typename ST::T(1)) return dens(p);
    else return dens(p) * (typename ST::T(1) / ST::kUnit);
}

// Momentum with the reference's Cx / Cy sums; the zero terms (f*0) are
// dropped, which leaves every nonzero result unchanged.
template <class T>
__device__ inline void momentum(const T (&p)[Q], T& jx, T& jy) {
    jx = ((((((T(0) + p[1]) - p[3]) + p[5]) - p[6]) - p[7]) + p[8]);
    jy = ((((((T(0) + p[2]) - p[4]) + p[5]) + p[6]) - p[7]) - p[8]);
}
__device__ inline void momentum(const float (&p)[Q], float& jx, float& jy) {
    jx = ((((p[1] - p[3]) + p[5]) - p[6]) - p[7]) + p[8];
    jy = ((((p[2] - p[4]) + p[5]) + p[6]) - p[7]) - p[8];
}

// Reference feq (src/fluid.cpp:254-276), evaluated as written.
__device__ inline void feq_ref(double rho, double ux, double uy, double (&e)[Q]) {
    const double r0 = (4.0 / 9.0) * rho, r1 = (1.0 / 9.0) * rho, r2 = (1.0 / 36.0) * rho;
    const double uxx = ux * ux, uyy = uy * uy;
    const double a5 = ux + uy, a6 = -ux + uy, a7 = -ux - uy, a8 = ux - uy;
    const double usq = uxx + uyy;
    e[0] = r0 * (1.0 - 1.5 * usq);
    e[1] = r1 * (1.0 + 3.0 * ux + 4.5 * uxx - 1.5 * usq);
    e[2] = r1 * (1.0 + 3.0 * uy + 4.5 * uyy - 1.5 * usq);
    e[3] = r1 * (1.0 - 3.0 * ux + 4.5 * uxx - 1.5 * usq);
    e[4] = r1 * (1.0 - 3.0 * uy + 4.5 * uyy - 1.5 * usq);
    e[5] = r2 * (1.0 + 3.0 * a5 + 4.5 * a5 * a5 - 1.5 * usq);
    e[6] = r2 * (1.0 + 3.0 * a6 + 4.5 * a6 * a6 - 1.5 * usq);
    e[7] = r2 * (1.0 + 3.0 * a7 + 4.5 * a7 * a7 - 1.5 * usq);
    e[8] = r2 * (1.0 + 3.0 * a8 + 4.5 * a8 * a8 - 1.5 * usq);
}

// BGK relaxation of one component in deviation form, FP32:
//   d_i' = d_i + omega (deq_i - d_i + F_i / omega),
//   deq_i = feq_i - w_i rho0 = w_i drho + w_i rho (3 cu + 4.5 cu^2 - 1.5 u^2),
// with drho = rho - rho0 so small perturbations keep their digits.  Opposite
// directions share cu^2 (their polynomials differ only in the sign of the
// 3 cu term).
//
// Which roundings the FP32 result carries decides how close it stays to FP64
// over long runs, because in a steady state a site's rounding repeats every
// step.  Four choices:
//  * the non-equilibrium part deq_i - d_i is formed directly:
//    fma(w_i, drho, -d_i) rounds once at the size of that (small) difference,
//    and the velocity terms are added to it, instead of rounding deq_i itself
//    at the size of the populations.  Rounded at that size, deq_i of
//    opposite directions carries independent errors, a frozen field of
//    spurious momentum sources; an FP32 emulation of a porous two-phase
//    lattice (512x256, 2000 steps, against FP64) ended 2.97e-5 from FP64 in
//    density that way and 1.97e-7 this way, the level of the FP32 storage
//    rounding alone;
//  * the body force enters before the one rounding of d_i' (as F_i / omega
//    inside the fma).  Added after it, the sum of a stored-grid value and the
//    fixed F_i rounds with the same bias at every site whose d_i shares a
//    binade, a systematic error in the force of up to half a unit in the last
//    place per step (the same emulation: 3e-5 from FP64 with the force added
//    after rounding, 3e-7 before);
//  * the update acts on the non-equilibrium part, which is small, rather than
//    adding (1 - omega) d_i and omega deq_i, which nearly cancel for omega
//    near 2 (a porous two-phase lattice at tau = 0.707 drifted 4x further
//    from FP64 with the latter form);
//  * the rest population closes the site's mass balance:
//    d_0' = d_0 + sum_{i>0} (d_i - d_i') + sum_i F_i, summing the changes,
//    which are small, rather than the populations.  Collision conserves
//    mass; evaluating d_0' from its own formula added (rho - rho0) times the
//    rounding of the weights to every site's mass at every step, always with
//    the same sign (1e-4 in density after 1000 steps of a 1.0 / 0.06 droplet
//    at tau = 0.8), and d_0' = drho - sum_{i>0} d_i' still drifted: in a
//    steady state the rounding of those large partial sums repeats every
//    step (a porous two-phase run lost 2e-5 of a component's mass in 1e4
//    steps; summing the changes cut an FP32 emulation's drift 1000x).
// force: the deviation forms' encoding of the body force (set_force):
// force[i] = F_i / omega for i > 0 and force[0] = sum_i F_i.
// ST: the storage form.  The mass balance sums the changes of the values as
// they will be stored (rounded to FP16 for StoreF16), so only the rounding of
// d_0' itself is left out of it: closed on the FP32 values, every FP16
// rounding of d_1'..d_8' was a mass error, and where a population changes by
// less than half its FP16 spacing the rounding has the same sign every step
// (a 4096^2 spinodal lattice lost 0.45% of its mass in 1000 steps that way).
// enc: the outputs as the storage form encodes them (their stores and the
// closure share the one encoding).
// p, out, r and the force terms are in working units (ST::kUnit times the
// deviations, see lattice.h); the weights carry the scale, exactly.
template <class ST>
__device__ inline void relax_dev(const float (&p)[Q], float rho, float drho, float ux, float uy, float omega,
                                 const float* force, bool has_force, float (&out)[Q], typename ST::S (&enc)[Q]) {
    constexpr float U = ST::kUnit;
    // Every fused multiply-add is written out (the file is built without
    // contraction): the compiler's own choice of which products to fuse
    // could differ between the kernels this is inlined into, and the
    // single-step and two-step kernels must give the same bits.
    const float usq15 = 1.5f * fmaf(ux, ux, uy * uy);
    const float wr1 = ((1.0f / 9.0f) * U) * rho, wr2 = ((1.0f / 36.0f) * U) * rho;
    float r = 0.f;   // sum over i > 0 of (d_i - d_i')
    auto pair = [&](int i, int j, float cu, float w, float wr) {
        const float t = fmaf(4.5f * cu, cu, -usq15);
        const float c3 = 3.0f * cu;
        // deq - d = (w drho - d) + w rho (t +- 3 cu)
        float ni = fmaf(wr, t + c3, fmaf(w, drho, -p[i]));
        float nj = fmaf(wr, t - c3, fmaf(w, drho, -p[j]));
        if (has_force) {
            ni += force[i] * U;
            nj += force[j] * U;
        }
        const typename ST::S ei = ST::encw(fmaf(omega, ni, p[i]));
        const typename ST::S ej = ST::encw(fmaf(omega, nj, p[j]));
        const float oi = ST::decw(ei);   // ST::rnd of the relaxed value, in working units
        const float oj = ST::decw(ej);
        out[i] = oi;
        out[j] = oj;
        enc[i] = ei;
        enc[j] = ej;
        r += (p[i] - oi) + (p[j] - oj);
    };
    pair(1, 3, ux, (1.0f / 9.0f) * U, wr1);
    pair(2, 4, uy, (1.0f / 9.0f) * U, wr1);
    pair(5, 7, ux + uy, (1.0f / 36.0f) * U, wr2);
    pair(6, 8, uy - ux, (1.0f / 36.0f) * U, wr2);
    out[0] = has_force ? (p[0] + r) + force[0] * U : p[0] + r;
    enc[0] = ST::encw(out[0]);
}

// Reciprocal for the deviation forms (hardware approximation, 1 ulp).
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ inline bool is_dry(double r) { return fabs(r) <= kZeroTol * fmax(1.0, fabs(r)); }
// The same test for an FP32 density, in FP32: for finite r it holds exactly
// when |r| <= kZeroTol (the factor max(1, |r|) only matters above 1, where
// the test fails), and for infinite r.  kZeroTolF is the largest float not
// above kZeroTol.
constexpr float kZeroTolF = 0x1.cd2b28p-54f;
static_assert((double)kZeroTolF <= kZeroTol && (double)0x1.cd2b2ap-54f > kZeroTol, "kZeroTolF");
__device__ inline bool is_dry(float r) {
    const float m = fabsf(r);
    return m <= kZeroTolF || m == __builtin_inff();
}

// The zero-density rule (src/lbm.cpp:262-269) keeps the previous step's
// equilibrium velocity where |rho| <= 1e-16.  The fused FP64 kernel holds it
// in HBM only where rho < kUeqWatch (ueq_hold), so a site whose density drops
// from >= kUeqWatch to <= 1e-16 in one step would reuse an older value.  Each
// hold records its step in stamp[s]; a reuse at step n (ueq_reuse) checks
// that the value held is step n-1's (or that the host declared every value
// valid since, stamp[plane + 1]) and otherwise raises the flag stamp[plane],
// which the host turns into an error (SCL_ERR_INEXACT).  The deviation forms,
// which are not bit-exact with the reference anyway, hold the velocity below
// thresholds far above their density noise (StoreF32 / StoreF16 ::kUeqWatch)
// and keep no records.
__device__ inline void ueq_hold(unsigned* stamp, long long s_idx, unsigned n) { G(stamp)[s_idx] = n; }
__device__ inline void ueq_reuse(unsigned* stamp, long long plane, long long s_idx, unsigned n) {
    const unsigned last = G(stamp)[s_idx], valid = G(stamp)[plane + 1];
    if ((int)(n - 1u - last) > 0 && (int)(n - 1u - valid) > 0)
        __hip_atomic_fetch_or(stamp + plane, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ueq_hold(stamp, s_idx, n);
}

// Bounce-back mode of the gathering kernels: 0 selects the pull source per
// direction from the mask bits; 1 / 2 read every pull from its source slot
// and keep the bounce-back values in the solid sites' slots (ghost stores,
// streamed / cached).
#ifndef SCL_GHOST
#define SCL_GHOST 0
#endif

// Byte offsets of rows y-1, y, y+1 and columns x-1, x, x+1 around a site
// (periodic), for the ghost stores.
template <class OffT>
struct Nbr {
    OffT rs, rc, rn, cm, cc, cp;
};

// Ghost stores.  A fluid site x whose neighbour x + c_i is solid also writes
// its post-collision population i into slot opp(i) of that neighbour.  The
// next step's pull of direction opp(i) at x reads exactly that slot, and the
// value there is the link bounce-back value P_i(x), so the gather needs no
// per-direction select and its loads do not wait for mask bits.  Slot j of a
// solid site s is written only by the fluid site s + c_j, so there is one
// writer per slot.  Solid sites are never read as sources of their own.
template <class ST, class OffT>
__device__ inline void ghost_stores(void* dst, OffT plane_b, const Nbr<OffT>& n, unsigned nb,
                                    const typename ST::T (&v)[Q]) {
    if (nb == 0x1FFu) return;   // no solid neighbour
#pragma unroll
    for (int i = 1; i < Q; ++i) {
        if (!((nb >> i) & 1u)) {
            const OffT row = cy(i) == 1 ? n.rn : (cy(i) == -1 ? n.rs : n.rc);
            const OffT col = cx(i) == 1 ? n.cp : (cx(i) == -1 ? n.cm : n.cc);
            typename ST::S* q = (typename ST::S*)((char*)dst + ((OffT)opp(i) * plane_b + row + col));
            if (SCL_GHOST == 2) ST::sce(q, ST::encw(v[i]));
            else ST::se(q, ST::encw(v[i]));
        }
    }
}

// ---------------------------------------------------------------------------
// Per-site completion: everything after the gather, for one site whose
// (gathered) populations are p, densities rho, neighbourhood bits nb and
// neighbour densities nr[c][i] (i = 1..8, valid where bit i of nb is set).
// ---------------------------------------------------------------------------
template <class ST, bool REPLAY, class OffT, bool GW>
__device__ inline void finish1(const CompArgs<typename ST::T>& ca, void* dst, long long plane, OffT plane_b,
                               int s_idx, OffT home, const typename ST::T (&p)[Q], typename ST::T rho_in, bool fl,
                               unsigned nb, const Nbr<OffT>& nbr) {
    if constexpr (!ST::kDev) {
        double ux = 0.0, uy = 0.0;
        const double rho = rho_in;
        if (fl) {
            double jx, jy;
            momentum(p, jx, jy);
            ux = jx / rho;   // src/fluid.cpp:196-197
            uy = jy / rho;
        }
        double e[Q];
        feq_ref(rho, ux, uy, e);
        if (REPLAY) {
            G(ca.rho)[s_idx] = rho;
            G(ca.jx)[s_idx] = ux;
            G(ca.jy)[s_idx] = uy;
#pragma unroll
            for (int i = 0; i < Q; ++i) G(ca.feq)[i * plane + s_idx] = e[i];
        } else if (fl) {
            double v[Q];
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                v[i] = ca.omc * p[i] + ca.omega * e[i];
                if (ca.has_force) v[i] += ca.force[i];
                st_at<ST>(dst, (OffT)i * plane_b + home, v[i]);
            }
            if constexpr (GW) ghost_stores<ST, OffT>(dst, plane_b, nbr, nb, v);
        }
    } else {
        if (!REPLAY && fl) {
            const float drho = rho_in;
            const float rho = ca.rho0 + drho;
            float jx, jy;
            momentum(p, jx, jy);   // working units
            const float inv = rcp(rho);
            constexpr float iu = 1.0f / ST::kUnit;
            float out[Q];
            typename ST::S enc[Q];
            relax_dev<ST>(p, rho, drho, (jx * inv) * iu, (jy * inv) * iu, ca.omega, ca.force, ca.has_force, out, enc);
#pragma unroll
            for (int i = 0; i < Q; ++i) ST::se((typename ST::S*)((char*)dst + ((OffT)i * plane_b + home)), enc[i]);
            if constexpr (GW) ghost_stores<ST, OffT>(dst, plane_b, nbr, nb, out);
        }
    }
}

// Two-component collision of one fluid site in the deviation forms (FP32
// arithmetic): fluid-fluid and fluid-solid forces from the neighbour
// densities nr (valid where bit i of nb is set), the composite and
// equilibrium velocities, and the relaxation.  Zero-density rule: a dry
// component takes the equilibrium velocity of the previous step from
// ueqx_in / ueqy_in and carries it forward into ueqx / ueqy; a component
// below kUeqWatch records its new one there.
// Equilibrium-velocity buffers of one step: [component] read (previous
// step's) and written.
struct UeqPtrs {
    const double* in_x[2];
    const double* in_y[2];
    double* out_x[2];
    double* out_y[2];
    __device__ const double* ix(int c) const { return in_x[c]; }
    __device__ const double* iy(int c) const { return in_y[c]; }
    __device__ double* ox(int c) const { return out_x[c]; }
    __device__ double* oy(int c) const { return out_y[c]; }
};

template <class T>
__device__ inline UeqPtrs ueq_ptrs(const StepArgs<T>& a) {
    return UeqPtrs{{a.c[0].ueqx_in, a.c[1].ueqx_in}, {a.c[0].ueqy_in, a.c[1].ueqy_in},
                   {a.c[0].ueqx, a.c[1].ueqx}, {a.c[0].ueqy, a.c[1].ueqy}};
}

// The collision with the physics parameters of `a` and the equilibrium-
// velocity buffers of `u` (the multi-step kernels use one parameter set for
// all of their steps).  U: UeqPtrs, or any type with the same accessors
// ix / iy (read, component c) and ox / oy (written).
// FORCE = false: no body force on either component (has_force is then not
// read, so kernels built for force-free lattices carry neither the force
// values nor the additions).
template <class T, class U, bool FORCE = true, class ST = StoreF32, bool OPEN = false>
__device__ inline void collide2_core(const StepArgs<T>& a, const U& u, int s_idx, const float (&p)[2][Q],
                                     const float (&rc)[2], const float (&nr)[2][Q], unsigned nb, float (&oa)[Q],
                                     float (&ob)[Q], typename ST::S (&ea)[Q], typename ST::S (&eb)[Q]);

template <class T, class ST, bool OPEN = false>
__device__ inline void collide2_dev(const StepArgs<T>& a, int s_idx, const float (&p)[2][Q], const float (&rc)[2],
                                    const float (&nr)[2][Q], unsigned nb, float (&oa)[Q], float (&ob)[Q],
                                    typename ST::S (&ea)[Q], typename ST::S (&eb)[Q]) {
    collide2_core<T, UeqPtrs, true, ST, OPEN>(a, ueq_ptrs(a), s_idx, p, rc, nr, nb, oa, ob, ea, eb);
}

template <class T, class U, bool FORCE, class ST, bool OPEN>
__device__ inline void collide2_core(const StepArgs<T>& a, const U& u, int s_idx, const float (&p)[2][Q],
                                     const float (&rc)[2], const float (&nr)[2][Q], unsigned nb, float (&oa)[Q],
                                     float (&ob)[Q], typename ST::S (&ea)[Q], typename ST::S (&eb)[Q]) {
    // fused multiply-adds written out, as in relax_dev
    const auto& A = a.c[0];
    const auto& B = a.c[1];
    // Neighbour sums sum_i w_i c_i v(x + c_i) as differences of opposite
    // neighbours first:
    //   x: (v1 - v3)/9 + ((v5 - v7) + (v8 - v6))/36,
    //   y: (v2 - v4)/9 + ((v5 - v7) - (v8 - v6))/36.
    // Summed term by term, the rounding is relative to terms much larger than
    // the result, and in a steady state it repeats every step and adds up to
    // a net force (a static droplet drifted 1.4e-2 lattice units in 2e4
    // steps); the differences of neighbouring values are small and nearly
    // exact.  Solid neighbours contribute zero.
    float v[2][Q], fl[Q];
#pragma unroll
    for (int i = 1; i < Q; ++i) {
        const bool f = (nb >> i) & 1u;
        v[0][i] = f ? nr[0][i] : 0.f;
        v[1][i] = f ? nr[1][i] : 0.f;
        fl[i] = f ? 1.f : 0.f;
    }
    auto nsum = [](const float (&a)[Q], float& sx, float& sy) {
        const float d57 = a[5] - a[7], d86 = a[8] - a[6];
        sx = fmaf(1.0f / 9.0f, a[1] - a[3], (1.0f / 36.0f) * (d57 + d86));
        sy = fmaf(1.0f / 9.0f, a[2] - a[4], (1.0f / 36.0f) * (d57 - d86));
    };
    float sd[2][2];   // [c]: sums over fluid nbrs of w c drho of the other component
    float sfx, sfy;   // sums over fluid nbrs of w c
    nsum(v[1], sd[0][0], sd[0][1]);
    nsum(v[0], sd[1][0], sd[1][1]);
    nsum(fl, sfx, sfy);
    const float da0 = rc[0], db0 = rc[1];
    const float ra = A.rho0 + da0, rb = B.rho0 + db0;
    // psi_other over fluid neighbours = sum w c drho + rho0_other * sum w c
    // OPEN (the kernels of lattices without solids): sum w c over the fluid
    // neighbours is +0, so the rho0 terms and the adhesion drop out.
    // fma(a, b, +-0) and a * b differ only in the sign of a zero result,
    // which the next sum with a nonzero term erases, so both forms give the
    // same populations.
    constexpr bool open = OPEN;
    const float ga = A.gneg * ra, gb = B.gneg * rb;
    const float fa_x = ga * (open ? sd[0][0] : fmaf(B.rho0, sfx, sd[0][0]));
    const float fa_y = ga * (open ? sd[0][1] : fmaf(B.rho0, sfy, sd[0][1]));
    const float fb_x = gb * (open ? sd[1][0] : fmaf(A.rho0, sfx, sd[1][0]));
    const float fb_y = gb * (open ? sd[1][1] : fmaf(A.rho0, sfy, sd[1][1]));
    // adhesion: sum over solid nbrs of w c = -(sum over fluid nbrs)
    const float fs_ax = A.gsneg * -sfx, fs_ay = A.gsneg * -sfy;
    const float fs_bx = B.gsneg * -sfx, fs_by = B.gsneg * -sfy;
    float jax, jay, jbx, jby;
    momentum(p[0], jax, jay);
    momentum(p[1], jbx, jby);
    const float inv_den = rcp(fmaf(ra, A.omega, rb * B.omega));
    constexpr float iu = 1.0f / ST::kUnit;   // the momenta are in working units
    const float ucx = (fmaf(jax, A.omega, jbx * B.omega) * inv_den) * iu;
    const float ucy = (fmaf(jay, A.omega, jby * B.omega) * inv_den) * iu;
    // Equilibrium velocities; at dry sites (|rho| <= 1e-16) the held ones
    // replace them below.
    float uea[2], ueb[2];
    {
        const float ir = rcp(ra);
        uea[0] = fmaf(A.inv_omega, open ? fa_x * ir : fmaf(fa_x, ir, fs_ax), ucx);
        uea[1] = fmaf(A.inv_omega, open ? fa_y * ir : fmaf(fa_y, ir, fs_ay), ucy);
    }
    {
        const float ir = rcp(rb);
        ueb[0] = fmaf(B.inv_omega, open ? fb_x * ir : fmaf(fb_x, ir, fs_bx), ucx);
        ueb[1] = fmaf(B.inv_omega, open ? fb_y * ir : fmaf(fb_y, ir, fs_by), ucy);
    }
    // The zero-density rule, behind one wave-uniform branch: sites below the
    // watch threshold are rare, and per-lane branches for each component cost
    // the common path their compares, exec-mask updates and the buffer
    // addresses.
    const bool dry_a = is_dry(ra), dry_b = is_dry(rb);
    const bool held_a = dry_a || fabsf(ra) < ST::kUeqWatch, held_b = dry_b || fabsf(rb) < ST::kUeqWatch;
    if (__builtin_amdgcn_ballot_w64(held_a || held_b) != 0) {
        if (dry_a) {
            const double ux0 = G(u.ix(0))[s_idx], uy0 = G(u.iy(0))[s_idx];
            uea[0] = (float)ux0;
            uea[1] = (float)uy0;
            G(u.ox(0))[s_idx] = ux0;
            G(u.oy(0))[s_idx] = uy0;
        } else if (held_a) {
            G(u.ox(0))[s_idx] = uea[0];
            G(u.oy(0))[s_idx] = uea[1];
        }
        if (dry_b) {
            const double ux0 = G(u.ix(1))[s_idx], uy0 = G(u.iy(1))[s_idx];
            ueb[0] = (float)ux0;
            ueb[1] = (float)uy0;
            G(u.ox(1))[s_idx] = ux0;
            G(u.oy(1))[s_idx] = uy0;
        } else if (held_b) {
            G(u.ox(1))[s_idx] = ueb[0];
            G(u.oy(1))[s_idx] = ueb[1];
        }
    }
    relax_dev<ST>(p[0], ra, da0, uea[0], uea[1], A.omega, A.force, FORCE && A.has_force, oa, ea);
    relax_dev<ST>(p[1], rb, db0, ueb[0], ueb[1], B.omega, B.force, FORCE && B.has_force, ob, eb);
}

// The reference's two-component arithmetic in FP64, operation by operation
// (src/lbm.cpp:110-140, 163-204, 250-269, 308-309, 335-336), in pieces that
// the single-step kernel (finish2) and the two-step FP64 tile share, so that
// both evaluate the same expressions.
//
// Fluid-fluid repulsion and fluid-solid adhesion of a fluid site: rep[c] =
// (sum over fluid neighbours of (w_i c_i) psi_other) * (G_neg,c psi_c),
// fs[c] = G_sneg,c * (sum over solid neighbours of w_i c_i).  A skipped
// (solid) neighbour adds +0.0, which leaves the sum unchanged: it starts at
// +0.0 and a round-to-nearest sum is -0.0 only when both terms are, so
// selects replace the reference's `continue` exactly.
__device__ inline void forces2_ref(const StepArgs<double>& a, const double (&rc)[2], const double (&nr)[2][Q],
                                   unsigned nb, double (&rep)[2][2], double (&fs)[2][2]) {
    const auto& A = a.c[0];
    const auto& B = a.c[1];
    const double ra = rc[0], rb = rc[1];
    double sax = 0.0, say = 0.0, sbx = 0.0, sby = 0.0;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
        const bool f = (nb >> i) & 1u;
        const double pa = nr[0][i], pb = nr[1][i];
        if (cx(i) != 0) {
            const double wc = wint(i) * cx(i);
            sax += f ? wc * pb : 0.0;
            sbx += f ? wc * pa : 0.0;
        }
        if (cy(i) != 0) {
            const double wc = wint(i) * cy(i);
            say += f ? wc * pb : 0.0;
            sby += f ? wc * pa : 0.0;
        }
    }
    rep[0][0] = sax * (A.gneg * ra);
    rep[0][1] = say * (A.gneg * ra);
    rep[1][0] = sbx * (B.gneg * rb);
    rep[1][1] = sby * (B.gneg * rb);
    double ssx = 0.0, ssy = 0.0;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
        const bool solid = !((nb >> i) & 1u);
        ssx += solid ? wint(i) * cx(i) : 0.0;
        ssy += solid ? wint(i) * cy(i) : 0.0;
    }
    fs[0][0] = A.gsneg * ssx;
    fs[0][1] = A.gsneg * ssy;
    fs[1][0] = B.gsneg * ssx;
    fs[1][1] = B.gsneg * ssy;
}
// Composite velocity u' = (j_A w_A + j_B w_B) / (rho_A w_A + rho_B w_B).
__device__ inline void composite_ref(const StepArgs<double>& a, const double (&jm)[2][2], double ra, double rb,
                                     double& ucx, double& ucy) {
    const auto& A = a.c[0];
    const auto& B = a.c[1];
    const double den = ra * A.omega + rb * B.omega;
    ucx = (jm[0][0] * A.omega + jm[1][0] * B.omega) / den;
    ucy = (jm[0][1] * A.omega + jm[1][1] * B.omega) / den;
}
// Equilibrium velocity of a non-dry component, one direction.
__device__ inline double ueq_ref(const CompArgs<double>& C, double uc, double rep, double fs, double r) {
    return uc + rep * C.inv_omega / r + fs * C.inv_omega;
}
// BGK relaxation and the volume force of one population.
__device__ inline double bgk_ref(const CompArgs<double>& C, int i, double p, double e, bool force) {
    double v = C.omc * p + C.omega * e;
    if (force) v += C.force[i];
    return v;
}

// The two-step FP64 tile's collision of one fluid site: the same
// expressions as finish2, with the equilibrium velocities taken only from
// the step's own state.  The zero-density rule's held velocities live in
// HBM for sites below kUeqWatch, and a tile's halo sites are computed by
// two workgroups, so where either component's density is below `watch`
// (>= kUeqWatch; or dry, or not finite) the site raises *flag (= launch)
// instead, and the host repeats the launch's steps one per launch (whose
// kernel keeps the records).  Elsewhere the result is bit for bit finish2's.
template <bool FORCE>
__device__ inline void collide2_f64(const StepArgs<double>& a, unsigned* flag, unsigned launch, double watch,
                                    const double (&p)[2][Q], const double (&rc)[2], const double (&nr)[2][Q],
                                    unsigned nb, double (&oa)[Q], double (&ob)[Q]) {
    double rep[2][2], fs[2][2], jm[2][2], ucx, ucy;
    forces2_ref(a, rc, nr, nb, rep, fs);
    momentum(p[0], jm[0][0], jm[0][1]);
    momentum(p[1], jm[1][0], jm[1][1]);
    composite_ref(a, jm, rc[0], rc[1], ucx, ucy);
    const double ma = fabs(rc[0]), mb = fabs(rc[1]);
    const bool plain = ma >= watch && ma <= __DBL_MAX__ && mb >= watch && mb <= __DBL_MAX__;
    if (!plain) __hip_atomic_fetch_or(flag, launch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const auto& C = a.c[c];
        double e[Q];
        feq_ref(rc[c], ueq_ref(C, ucx, rep[c][0], fs[c][0], rc[c]), ueq_ref(C, ucy, rep[c][1], fs[c][1], rc[c]), e);
        double (&o)[Q] = c == 0 ? oa : ob;
#pragma unroll
        for (int i = 0; i < Q; ++i) o[i] = bgk_ref(C, i, p[c][i], e[i], FORCE && C.has_force);
    }
}

template <class ST, bool REPLAY, class OffT, bool GW, bool OPEN = false>
__device__ inline void finish2(const StepArgs<typename ST::T>& a, void* const (&dst)[2], OffT plane_b, int s_idx,
                               OffT home,
                               const typename ST::T (&p)[2][Q], const typename ST::T (&rc)[2],
                               const typename ST::T (&nr)[2][Q], unsigned nb, const Nbr<OffT>& nbr) {
    const bool fl = nb & 1u;
    const long long plane = a.plane;
    if constexpr (!ST::kDev) {
        const auto& A = a.c[0];
        const auto& B = a.c[1];
        const double ra = rc[0], rb = rc[1];
        double rep[2][2] = {{0.0, 0.0}, {0.0, 0.0}};   // [comp][x/y]
        double fs[2][2];
        double jm[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
        double ue[2][2];
        double e[2][Q];
        if (fl) {
            forces2_ref(a, rc, nr, nb, rep, fs);
            momentum(p[0], jm[0][0], jm[0][1]);
            momentum(p[1], jm[1][0], jm[1][1]);
            double ucx, ucy;
            composite_ref(a, jm, ra, rb, ucx, ucy);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const auto& C = a.c[c];
                const double r = rc[c];
                if (!is_dry(r)) {
                    ue[c][0] = ueq_ref(C, ucx, rep[c][0], fs[c][0], r);
                    ue[c][1] = ueq_ref(C, ucy, rep[c][1], fs[c][1], r);
                    if (!REPLAY && fabs(r) < kUeqWatch) {
                        G(C.ueqx)[s_idx] = ue[c][0];
                        G(C.ueqy)[s_idx] = ue[c][1];
                        ueq_hold(C.stamp, s_idx, C.step);
                    }
                } else {
                    ue[c][0] = G(C.ueqx)[s_idx];
                    ue[c][1] = G(C.ueqy)[s_idx];
                    if (!REPLAY) ueq_reuse(C.stamp, plane, s_idx, C.step);
                }
            }
        } else if (REPLAY) {
            ue[0][0] = G(A.ueqx)[s_idx];
            ue[0][1] = G(A.ueqy)[s_idx];
            ue[1][0] = G(B.ueqx)[s_idx];
            ue[1][1] = G(B.ueqy)[s_idx];
        }
        if (fl || REPLAY) {
#pragma unroll
            for (int c = 0; c < 2; ++c) feq_ref(rc[c], ue[c][0], ue[c][1], e[c]);
        }
        if (REPLAY) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const auto& C = a.c[c];
                G(C.rho)[s_idx] = rc[c];
                G(C.jx)[s_idx] = jm[c][0];
                G(C.jy)[s_idx] = jm[c][1];
                G(C.frx)[s_idx] = rep[c][0];
                G(C.fry)[s_idx] = rep[c][1];
                if (fl) {
                    G(C.ueqx)[s_idx] = ue[c][0];
                    G(C.ueqy)[s_idx] = ue[c][1];
                }
#pragma unroll
                for (int i = 0; i < Q; ++i) G(C.feq)[i * plane + s_idx] = e[c][i];
            }
        } else if (fl) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const auto& C = a.c[c];
                double v[Q];
#pragma unroll
                for (int i = 0; i < Q; ++i) {
                    v[i] = bgk_ref(C, i, p[c][i], e[c][i], C.has_force);
                    st_at<ST>(dst[c], (OffT)i * plane_b + home, v[i]);
                }
                if constexpr (GW) ghost_stores<ST, OffT>(dst[c], plane_b, nbr, nb, v);
            }
        }
    } else {
        // deviation forms, FP32 arithmetic
        if (!REPLAY && fl) {
            float oa[Q], ob[Q];
            typename ST::S ea[Q], eb[Q];
            collide2_dev<typename ST::T, ST, OPEN>(a, s_idx, p, rc, nr, nb, oa, ob, ea, eb);
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                ST::se((typename ST::S*)((char*)dst[0] + ((OffT)i * plane_b + home)), ea[i]);
                ST::se((typename ST::S*)((char*)dst[1] + ((OffT)i * plane_b + home)), eb[i]);
            }
            if constexpr (GW) {
                ghost_stores<ST, OffT>(dst[0], plane_b, nbr, nb, oa);
                ghost_stores<ST, OffT>(dst[1], plane_b, nbr, nb, ob);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// The fused kernel.  NC components, storage ST, NW wavefronts per workgroup,
// GATHER (false only for the first step after an upload: populations are the
// reference's streamed state and are read in place), REPLAY (recompute the
// step and write the reference's intermediate arrays instead of P_n), OffT
// the byte-offset type (32-bit while one component's buffer is < 4 GiB), LA
// the row lookahead of the software pipeline (1 or 2), OPEN a lattice with
// no solid site (the mask is not read and every neighbourhood is fluid, so
// the bounce-back selects and the adhesion sums fold away at compile time).
// ---------------------------------------------------------------------------
// Occupancy request: the solid-free FP32-arithmetic kernels fit in 168
// registers, i.e. three waves per SIMD (four measured no faster with FP32
// storage, and with the FP64 force sums they would spill); with FP16
// storage in 128, four waves per SIMD (left at three, the kernel took 138
// registers and ran 31% slower at 4096^2).
template <class ST, bool OPEN, int LA>
#ifndef SCL_GATHER_WAVES
#define SCL_GATHER_WAVES 1
#endif
constexpr int step_waves_per_eu() {
    return (OPEN && LA == 1 && sizeof(typename ST::T) == 4)
               ? (sizeof(typename ST::S) == 2 ? 4 : 3)
               : ((!OPEN && sizeof(typename ST::T) == 4) ? SCL_GATHER_WAVES : 1);
}

// Source and destination population buffers of one launch, per component.
struct Bufs {
    const void* src[2];
    void* dst[2];
};

// LDS of one workgroup: the density ring (4 rows, both components, strip
// plus one halo column at each side).
template <int NC, class T, int NW>
using DensRing = T[NC == 2 ? 4 : 1][NC][64 * NW + 2];

// One tile of the fused step: the strip of columns [64*NW*bx, 64*NW*(bx+1))
// over rows [y0, y1), for the lattice and parameters in `a`, reading the
// populations in b.src and writing b.dst.  Every thread of the workgroup
// calls it with the same arguments (it contains workgroup barriers).
template <int NC, class ST, int NW, bool GATHER, bool REPLAY, class OffT, int LA, bool OPEN>
__device__ __forceinline__ void step_tile(const StepArgs<typename ST::T>& a, const Bufs& b, int bx, int y0, int y1,
                                          DensRing<NC, typename ST::T, NW>& lds) {
    using T = typename ST::T;
    using S = typename ST::S;
    constexpr int BW = 64 * NW;
    constexpr OffT ES = sizeof(S);

    const int nx = a.nx, ny = a.ny, wpr = a.wpr;
    const OffT plane_b = (OffT)a.plane * ES;
    const OffT pitch_b = (OffT)a.pitch * ES;
    const uint64_t* __restrict__ mask = a.mask;
    // fluid bits (west, centre, east) of one row around a column, and the
    // neighbourhood of a site from three of them (all fluid in OPEN)
    auto tri = [&](int row, int cm, int c, int cp) -> unsigned {
        if constexpr (OPEN) return 7u;
        else return triple(mask, wpr, row, cm, c, cp);
    };
    auto neigh = [&](unsigned s, unsigned c, unsigned n) -> unsigned {
        if constexpr (OPEN) return 0x1FFu;
        else return neigh_bits(s, c, n);
    };

    const int tid = threadIdx.x;
    const int x0 = bx * BW;
    const int x = x0 + tid;
    const bool xin = x < nx;
    const int xm = xin ? wrap1(x - 1, nx) : 0;
    const int xp = xin ? wrap1(x + 1, nx) : 0;

    // Halo columns (densities only).  The left halo belongs to thread 0, the
    // right one to the last thread of the strip, or to thread 1 when the
    // strip is one column wide (so no thread holds both).
    const int tlast = min(BW, nx - x0) - 1;
    const int hrt = tlast == 0 ? 1 : tlast;
    const bool hrole = NC == 2 && (tid == 0 || tid == hrt);
    const int hcol = tid == 0 ? wrap1(x0 - 1, nx) : wrap1(x0 + tlast + 1, nx);
    const int hslot = tid == 0 ? 0 : tlast + 2;
    const int hcm = wrap1(hcol - 1, nx), hcp = wrap1(hcol + 1, nx);

    // Lookahead: the row whose loads are issued while row y is finished is
    // y+LA.  With LA = 1 (two register sets) its densities are filed before
    // the barrier; with LA = 2 (three sets) after finishing row y, so the
    // loads stay in flight during the arithmetic of row y.
    T pA[NC][Q], pB[NC][Q], pC[NC][Q], ph[NC][Q];
    unsigned nbA = 0, nbB = 0, nbC = 0, hnb = 0;
    // mask triples (west, centre, east bits) of rows y+LA-1 .. y+LA+1 for
    // the main column (t*) and the halo column (h*), rotated every row
    unsigned t1 = 0, t2 = 0, t3 = 0, h1 = 0, h2 = 0, h3 = 0;

    // Loads of one site (all components) in one block; a solid site in a
    // gathering step reads nothing (the reference's streamed state is zero
    // there) and its lane is simply masked off for the whole block.
    auto issue = [&](int col, int colm, int colp, int r, int rm, int rp, unsigned nb, T (&p)[NC][Q]) {
        OffT off[Q];
        pull_offsets<ST, GATHER, OffT, !REPLAY && SCL_GHOST != 0>(plane_b, pitch_b, col, colm, colp, r, rm, rp, nb,
                                                                 off);
        if (!GATHER || (nb & 1u)) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
#pragma unroll
                for (int i = 0; i < Q; ++i) p[c][i] = ldw_at<ST, OffT>(b.src[c], off[i]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
#pragma unroll
                for (int i = 0; i < Q; ++i) p[c][i] = T(0);
            }
        }
    };
    auto file_rho = [&](const T (&p)[NC][Q], int slot, int col_slot) {
        if constexpr (NC == 2) {
#pragma unroll
            for (int c = 0; c < NC; ++c) lds[slot][c][col_slot] = dens_w<ST>(p[c]);
        }
    };

    // ---- prologue: rows y0-1 (densities), y0 .. y0+LA-1 -------------------
    {
        const int r_m2 = wrap1(wrap1(y0 - 1, ny) - 1, ny), r_m1 = wrap1(y0 - 1, ny);
        const int r_p1 = wrap1(y0 + 1, ny), r_p2 = wrap1(r_p1 + 1, ny), r_p3 = wrap1(r_p2 + 1, ny);
        unsigned tm2 = 0, tm1 = 0, t0 = 0, tp1 = 0, tp2 = 0, tp3 = 0;
        unsigned hm2 = 0, hm1 = 0, h0 = 0, hp1 = 0, hp2 = 0, hp3 = 0;
        if (xin) {
            tm2 = tri(r_m2, xm, x, xp);
            tm1 = tri(r_m1, xm, x, xp);
            t0 = tri(y0, xm, x, xp);
            tp1 = tri(r_p1, xm, x, xp);
            tp2 = tri(r_p2, xm, x, xp);
            if (LA == 2) tp3 = tri(r_p3, xm, x, xp);
        }
        if (hrole) {
            hm2 = tri(r_m2, hcm, hcol, hcp);
            hm1 = tri(r_m1, hcm, hcol, hcp);
            h0 = tri(y0, hcm, hcol, hcp);
            hp1 = tri(r_p1, hcm, hcol, hcp);
            hp2 = tri(r_p2, hcm, hcol, hcp);
            if (LA == 2) hp3 = tri(r_p3, hcm, hcol, hcp);
        }
        if constexpr (NC == 2) {
            if (xin) {
                issue(x, xm, xp, r_m1, r_m2, y0, neigh(tm2, tm1, t0), pC);
                file_rho(pC, 0, tid + 1);
            }
            if (hrole) {
                issue(hcol, hcm, hcp, r_m1, r_m2, y0, neigh(hm2, hm1, h0), ph);
                file_rho(ph, 0, hslot);
                issue(hcol, hcm, hcp, y0, r_m1, r_p1, neigh(hm1, h0, hp1), ph);
                file_rho(ph, 1, hslot);
                if (LA == 2) {
                    issue(hcol, hcm, hcp, r_p1, y0, r_p2, neigh(h0, hp1, hp2), ph);
                    file_rho(ph, 2, hslot);
                }
            }
        }
        if (xin) {
            nbA = neigh(tm1, t0, tp1);
            issue(x, xm, xp, y0, r_m1, r_p1, nbA, pA);
            file_rho(pA, 1, tid + 1);
            if (LA == 2 && (NC == 2 || y0 + 1 < y1)) {
                nbB = neigh(t0, tp1, tp2);
                issue(x, xm, xp, r_p1, y0, r_p2, nbB, pB);
                file_rho(pB, 2, tid + 1);
            }
        }
        if (LA == 1) {
            t1 = t0; t2 = tp1; t3 = tp2;
            h1 = h0; h2 = hp1; h3 = hp2;
        } else {
            t1 = tp1; t2 = tp2; t3 = tp3;
            h1 = hp1; h2 = hp2; h3 = hp3;
        }
    }

    // ---- one row: issue row y+LA, finish row y ----------------------------
    auto body = [&](int y, T (&p0)[NC][Q], unsigned nb0, T (&pn)[NC][Q], unsigned& nbn) {
        const int k = y - y0;
        const bool more = NC == 2 ? (y + LA <= y1) : (y + LA < y1);
        const int rA = wrap1(y + LA, ny);              // row issued now
        const int rAm = wrap1(rA - 1, ny), rAp = wrap1(rA + 1, ny);
        const int slotA = (k + LA + 1) & 3;
        if (more) {
            if (xin) {
                nbn = neigh(t1, t2, t3);
                issue(x, xm, xp, rA, rAm, rAp, nbn, pn);
            }
            if (hrole) {
                hnb = neigh(h1, h2, h3);
                issue(hcol, hcm, hcp, rA, rAm, rAp, hnb, ph);
            }
            // mask row of the next issue
            const int rnext = wrap1(rAp + 1, ny);
            if (xin) {
                t1 = t2;
                t2 = t3;
                t3 = tri(rnext, xm, x, xp);
            }
            if (hrole) {
                h1 = h2;
                h2 = h3;
                h3 = tri(rnext, hcm, hcol, hcp);
            }
            if (LA == 1) {
                if (xin) file_rho(pn, slotA, tid + 1);
                if (hrole) file_rho(ph, slotA, hslot);
            }
        }
        if constexpr (NC == 2) __syncthreads();
        if (xin) {
            const int s_idx = y * a.pitch + x;
            const OffT home = (OffT)y * pitch_b + (OffT)x * ES;
            constexpr bool GW = !OPEN && !REPLAY && SCL_GHOST != 0;   // ghost stores (see ghost_stores)
            Nbr<OffT> nbr{};
            if constexpr (GW) {
                nbr.rs = (OffT)wrap1(y - 1, ny) * pitch_b;
                nbr.rc = (OffT)y * pitch_b;
                nbr.rn = (OffT)wrap1(y + 1, ny) * pitch_b;
                nbr.cm = (OffT)xm * ES;
                nbr.cc = (OffT)x * ES;
                nbr.cp = (OffT)xp * ES;
            }
            if constexpr (NC == 1) {
                finish1<ST, REPLAY, OffT, GW>(a.c[0], b.dst[0], a.plane, plane_b, s_idx, home, p0[0], dens_w<ST>(p0[0]),
                                              OPEN || (nb0 & 1u), OPEN ? 0x1FFu : nb0, nbr);
            } else {
                const int sS = k & 3, sC = (k + 1) & 3, sN = (k + 2) & 3;
                T nr[2][Q], r0[2];
#pragma unroll
                for (int i = 1; i < Q; ++i) {
                    const int sl = cy(i) == 1 ? sN : (cy(i) == -1 ? sS : sC);
#pragma unroll
                    for (int c = 0; c < 2; ++c) nr[c][i] = lds[sl][c][tid + 1 + cx(i)];
                }
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    r0[c] = lds[sC][c][tid + 1];
                    nr[c][0] = r0[c];
                }
                finish2<ST, REPLAY, OffT, GW, OPEN>(a, b.dst, plane_b, s_idx, home, p0, r0, nr, OPEN ? 0x1FFu : nb0, nbr);
            }
        }
        if (LA == 2 && more) {
            if (xin) file_rho(pn, slotA, tid + 1);
            if (hrole) file_rho(ph, slotA, hslot);
        }
    };

    if constexpr (LA == 1) {
        for (int y = y0; y < y1; y += 2) {
            body(y, pA, nbA, pB, nbB);
            if (y + 1 < y1) body(y + 1, pB, nbB, pA, nbA);
        }
    } else {
        for (int y = y0; y < y1; y += 3) {
            body(y, pA, nbA, pC, nbC);
            if (y + 1 < y1) body(y + 1, pB, nbB, pA, nbA);
            if (y + 2 < y1) body(y + 2, pC, nbC, pB, nbB);
        }
    }
}

// The fused kernel over one lattice: one workgroup per tile.  Workgroups are
// dealt round-robin to the 8 XCDs; consecutive ids of one XCD are given
// adjacent strips so the halo rows they share stay in that XCD's L2.
template <int NC, class ST, int NW, bool GATHER, bool REPLAY, class OffT, int LA, bool OPEN>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(step_waves_per_eu<ST, OPEN, LA>()))) void
k_step(const StepArgs<typename ST::T> a) {
    __shared__ DensRing<NC, typename ST::T, NW> lds;
    const int nbx = a.gx;
    const int nby = (a.ny + a.rows - 1) / a.rows;
    int bid = blockIdx.x;
    {
        const int nblk = nbx * nby;
        const int xcd = bid & 7, k = bid >> 3;
        const int per = (nblk + 7) >> 3;          // ids per XCD (the last ones may hold one fewer)
        const int full = nblk - 8 * (per - 1);    // XCDs holding `per` ids
        bid = xcd < full ? xcd * per + k : full * per + (xcd - full) * (per - 1) + k;
    }
    const int bx = bid % nbx, by = bid / nbx;
    Bufs b;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        b.src[c] = a.c[c].src;
        b.dst[c] = a.c[c].dst;
    }
    const int y0 = by * a.rows;
    step_tile<NC, ST, NW, GATHER, REPLAY, OffT, LA, OPEN>(a, b, bx, y0, min(y0 + a.rows, a.ny), lds);
}

}  // namespace scl
