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
// column.  Per row it gathers the populations of the row ahead (y+1), sums
// the densities of both components into a 4-slot LDS ring (with one halo
// column at each side of the strip, computed by the edge threads), waits at
// one barrier, and then finishes row y: the 8-neighbour Shan-Chen forces
// read the three LDS rows, the rest stays in registers.  The halo densities
// of the strip are recomputed from gathered populations, so no density
// array crosses HBM.
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
    double* ueqx;      // equilibrium velocity store (dry-site fallback / replay output)
    double* ueqy;
    // replay outputs (FP64, pitched), written only by the replay variant
    double* rho;
    double* jx;
    double* jy;
    double* frx;
    double* fry;
    double* feq;       // 9 planes
};

template <class T>
struct StepArgs {
    int nx, ny, pitch, wpr;
    long long plane;
    const uint64_t* mask;
    int rows;          // rows per workgroup
    int gx;            // workgroups along x
    CompArgs<T> c[2];
};

// Gather (or read in place) the 9 populations of component c at site
// (x, r); nb holds the fluid bits of the neighbourhood of (x, r).
template <class ST, bool GATHER>
__device__ inline void load_site(const typename ST::S* __restrict__ src, long long plane, int pitch, int x,
                                 int xm, int xp, int r, int rm, int rp, unsigned nb,
                                 typename ST::T (&p)[Q]) {
    const int o_r = r * pitch, o_s = rm * pitch, o_n = rp * pitch;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const typename ST::S* ptr;
        if (GATHER) {
            const int row = cy(i) == 1 ? o_s : (cy(i) == -1 ? o_n : o_r);
            const int col = cx(i) == 1 ? xm : (cx(i) == -1 ? xp : x);
            const bool ok = (nb >> opp(i)) & 1u;
            ptr = ok ? (src + i * plane + row + col) : (src + opp(i) * plane + o_r + x);
        } else {
            ptr = src + i * plane + o_r + x;
        }
        p[i] = ST::ld(ptr);
    }
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

// Momentum with the reference's Cx / Cy sums; the zero terms (f*0) are
// dropped, which leaves every nonzero result unchanged.
template <class T>
__device__ inline void momentum(const T (&p)[Q], T& jx, T& jy) {
    jx = ((((((T(0) + p[1]) - p[3]) + p[5]) - p[6]) - p[7]) + p[8]);
    jy = ((((((T(0) + p[2]) - p[4]) + p[5]) + p[6]) - p[7]) - p[8]);
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

// Equilibrium deviation e_i = feq_i - w_i*rho0 written with the density
// deviation drho = rho - rho0 so that small perturbations keep their digits:
// e_i = w_i*drho + w_i*rho*(3 cu + 4.5 cu^2 - 1.5 u^2).
__device__ inline void feq_dev(float rho, float drho, float ux, float uy, float (&e)[Q]) {
    const float usq15 = 1.5f * (ux * ux + uy * uy);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const float w = (float)weq(i);
        const float cu = (float)cx(i) * ux + (float)cy(i) * uy;
        const float poly = cu * (3.0f + 4.5f * cu) - usq15;
        e[i] = w * drho + (w * rho) * poly;
    }
}

__device__ inline bool is_dry(double r) { return fabs(r) <= kZeroTol * fmax(1.0, fabs(r)); }

// ---------------------------------------------------------------------------
// The fused kernel.  NC components, storage ST, NW wavefronts per workgroup,
// GATHER (false only for the first step after an upload: populations are the
// reference's streamed state and are read in place), REPLAY (recompute the
// step and write the reference's intermediate arrays instead of P_n).
// ---------------------------------------------------------------------------
template <int NC, class ST, int NW, bool GATHER, bool REPLAY>
__global__ __launch_bounds__(64 * NW) void k_step(const StepArgs<typename ST::T> a) {
    using T = typename ST::T;
    using S = typename ST::S;
    constexpr int BW = 64 * NW;
    constexpr bool DEV = ST::kDev;

    __shared__ T lds[NC == 2 ? 4 : 1][NC][BW + 2];

    const int nx = a.nx, ny = a.ny, pitch = a.pitch, wpr = a.wpr;
    const long long plane = a.plane;
    const uint64_t* __restrict__ mask = a.mask;

    // Workgroup -> strip.  Workgroups are dealt round-robin to the 8 XCDs;
    // consecutive ids of one XCD are given vertically adjacent strips so the
    // halo rows they share stay in that XCD's L2.
    const int nbx = a.gx;
    const int nby = (ny + a.rows - 1) / a.rows;
    int bid = blockIdx.x;
    {
        const int nblk = nbx * nby;
        const int xcd = bid & 7, k = bid >> 3;
        const int per = (nblk + 7) >> 3;      // ids per XCD (last XCD may hold fewer)
        const int full = nblk - 8 * (per - 1);  // XCDs holding `per` ids
        int lin;
        if (xcd < full) lin = xcd * per + k;
        else lin = full * per + (xcd - full) * (per - 1) + k;
        bid = lin;
    }
    const int bx = bid % nbx, by = bid / nbx;

    const int tid = threadIdx.x;
    const int x0 = bx * BW;
    const int x = x0 + tid;
    const bool xin = x < nx;
    const int xm = xin ? wrap1(x - 1, nx) : 0;
    const int xp = xin ? wrap1(x + 1, nx) : 0;
    const int y0 = by * a.rows;
    const int y1 = min(y0 + a.rows, ny);

    // halo columns (density only)
    const int tlast = min(BW, nx - x0) - 1;
    const bool lhalo = NC == 2 && tid == 0;
    const bool rhalo = NC == 2 && tid == tlast;
    const int hl = wrap1(x0 - 1, nx);
    const int hr = wrap1(x0 + tlast + 1, nx);
    const int hlm = wrap1(hl - 1, nx), hlp = wrap1(hl + 1, nx);
    const int hrm = wrap1(hr - 1, nx), hrp = wrap1(hr + 1, nx);

    const S* src[NC];
    S* dst[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        src[c] = (const S*)a.c[c].src;
        dst[c] = (S*)a.c[c].dst;
    }

    // Gather one site for all components, return the neighbourhood bits and
    // fill populations p and densities (or density deviations) rho.
    auto site = [&](int col, int colm, int colp, int r, int rm, int rp, unsigned ts, unsigned tc,
                    unsigned tn, T (&p)[NC][Q], T (&rho)[NC]) -> unsigned {
        const unsigned nb = neigh_bits(ts, tc, tn);
        const bool fl = nb & 1u;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (!GATHER || fl) {
                load_site<ST, GATHER>(src[c], plane, pitch, col, colm, colp, r, rm, rp, nb, p[c]);
            } else {
#pragma unroll
                for (int i = 0; i < Q; ++i) p[c][i] = T(0);
            }
            rho[c] = dens(p[c]);
        }
        return nb;
    };

    T pc[NC][Q], pn[NC][Q];
    T rc[NC], rn[NC];
    unsigned nbc = 0, nbn = 0;

    // rows needed: mask triples of y0-2 .. y1+1, gathered rows y0-1 .. y1
    const int rS2 = wrap1(wrap1(y0 - 1, ny) - 1, ny);
    const int rS1 = wrap1(y0 - 1, ny);
    const int rN1 = wrap1(y0 + 1, ny);
    unsigned t_s2 = 0, t_s1 = 0, t_0 = 0, t_1 = 0;
    if (xin) {
        t_s2 = triple(mask, wpr, rS2, xm, x, xp);
        t_s1 = triple(mask, wpr, rS1, xm, x, xp);
        t_0 = triple(mask, wpr, y0, xm, x, xp);
        t_1 = triple(mask, wpr, rN1, xm, x, xp);
    }

    if constexpr (NC == 2) {
        // row y0-1: densities only
        if (xin) {
            T pt[NC][Q], rt[NC];
            site(x, xm, xp, rS1, rS2, y0, t_s2, t_s1, t_0, pt, rt);
#pragma unroll
            for (int c = 0; c < NC; ++c) lds[0][c][tid + 1] = rt[c];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (!(h == 0 ? lhalo : rhalo)) continue;
            const int hc = h == 0 ? hl : hr, hm = h == 0 ? hlm : hrm, hp = h == 0 ? hlp : hrp;
            const int slot = h == 0 ? 0 : tid + 2;
#pragma unroll
            for (int k = 0; k < 2; ++k) {   // rows y0-1 and y0
                const int r = k == 0 ? rS1 : y0;
                const int rm = k == 0 ? rS2 : rS1;
                const int rp = k == 0 ? y0 : rN1;
                const unsigned ts = triple(mask, wpr, rm, hm, hc, hp);
                const unsigned tc = triple(mask, wpr, r, hm, hc, hp);
                const unsigned tn = triple(mask, wpr, rp, hm, hc, hp);
                T pt[NC][Q], rt[NC];
                site(hc, hm, hp, r, rm, rp, ts, tc, tn, pt, rt);
#pragma unroll
                for (int c = 0; c < NC; ++c) lds[k][c][slot] = rt[c];
            }
        }
    }
    if (xin) {
        nbc = site(x, xm, xp, y0, rS1, rN1, t_s1, t_0, t_1, pc, rc);
        if constexpr (NC == 2) {
#pragma unroll
            for (int c = 0; c < NC; ++c) lds[1][c][tid + 1] = rc[c];
        }
    }

    for (int y = y0; y < y1; ++y) {
        const int k = y - y0;
        const int ym = wrap1(y - 1, ny);
        const int yp = wrap1(y + 1, ny);
        const int ypp = wrap1(yp + 1, ny);
        // gather row y+1 (its populations are the next row's input)
        const bool more = (y + 1 < y1);
        if (NC == 2 || more) {
            if (xin) {
                const unsigned t_2 = triple(mask, wpr, ypp, xm, x, xp);
                nbn = site(x, xm, xp, yp, y, ypp, t_0, t_1, t_2, pn, rn);
                t_0 = t_1;
                t_1 = t_2;
                if constexpr (NC == 2) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) lds[(k + 2) & 3][c][tid + 1] = rn[c];
                }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (!(h == 0 ? lhalo : rhalo)) continue;
                const int hc = h == 0 ? hl : hr, hm = h == 0 ? hlm : hrm, hp = h == 0 ? hlp : hrp;
                const int slot = h == 0 ? 0 : tid + 2;
                const unsigned ts = triple(mask, wpr, y, hm, hc, hp);
                const unsigned tc = triple(mask, wpr, yp, hm, hc, hp);
                const unsigned tn = triple(mask, wpr, ypp, hm, hc, hp);
                T pt[NC][Q], rt[NC];
                site(hc, hm, hp, yp, y, ypp, ts, tc, tn, pt, rt);
#pragma unroll
                for (int c = 0; c < NC; ++c) lds[(k + 2) & 3][c][slot] = rt[c];
            }
        }
        if (NC == 2) __syncthreads();

        if (xin) {
            const int s_idx = y * pitch + x;
            const bool fl = nbc & 1u;
            if constexpr (NC == 1) {
                // ---------------- one component: collide(geom,f) + force -----------
                const auto& ca = a.c[0];
                if constexpr (!DEV) {
                    double ux = 0.0, uy = 0.0;
                    const double rho = rc[0];
                    if (fl) {
                        double jx, jy;
                        momentum(pc[0], jx, jy);
                        ux = jx / rho;   // src/fluid.cpp:196-197
                        uy = jy / rho;
                    }
                    double e[Q];
                    feq_ref(rho, ux, uy, e);
                    if (REPLAY) {
                        ca.rho[s_idx] = rho;
                        ca.jx[s_idx] = ux;
                        ca.jy[s_idx] = uy;
#pragma unroll
                        for (int i = 0; i < Q; ++i) ca.feq[i * plane + s_idx] = e[i];
                    } else if (fl) {
                        S* d = dst[0];
#pragma unroll
                        for (int i = 0; i < Q; ++i) {
                            double v = ca.omc * pc[0][i] + ca.omega * e[i];
                            if (ca.has_force) v += ca.force[i];
                            ST::st(d + i * plane + s_idx, v);
                        }
                    }
                } else {
                    if (!REPLAY && fl) {
                        const float drho = rc[0];
                        const float rho = ca.rho0 + drho;
                        float jx, jy;
                        momentum(pc[0], jx, jy);
                        const float ux = jx / rho, uy = jy / rho;
                        float e[Q];
                        feq_dev(rho, drho, ux, uy, e);
                        S* d = dst[0];
#pragma unroll
                        for (int i = 0; i < Q; ++i) {
                            float v = ca.omc * pc[0][i] + ca.omega * e[i];
                            if (ca.has_force) v += ca.force[i];
                            ST::st(d + i * plane + s_idx, v);
                        }
                    }
                }
            } else {
                // ---------------- two components: Shan-Chen ----------------------
                const int sS = k & 3, sC = (k + 1) & 3, sN = (k + 2) & 3;
                const int cl = tid + 1;
                if constexpr (!DEV) {
                    const auto& A = a.c[0];
                    const auto& B = a.c[1];
                    const double ra = rc[0], rb = rc[1];
                    double rep[2][2] = {{0.0, 0.0}, {0.0, 0.0}};   // [comp][x/y]
                    double jm[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
                    double ue[2][2];
                    double e[2][Q];
                    if (fl) {
                        // repulsion: sum over fluid neighbours of (w_i c_i) psi_other
                        double sax = 0.0, say = 0.0, sbx = 0.0, sby = 0.0;
#pragma unroll
                        for (int i = 1; i < Q; ++i) {
                            if ((nbc >> i) & 1u) {
                                const int sl = cy(i) == 1 ? sN : (cy(i) == -1 ? sS : sC);
                                const double pa = lds[sl][0][cl + cx(i)];
                                const double pb = lds[sl][1][cl + cx(i)];
                                if (cx(i) != 0) {
                                    const double wc = wint(i) * cx(i);
                                    sax += wc * pb;
                                    sbx += wc * pa;
                                }
                                if (cy(i) != 0) {
                                    const double wc = wint(i) * cy(i);
                                    say += wc * pb;
                                    sby += wc * pa;
                                }
                            }
                        }
                        rep[0][0] = sax * (A.gneg * ra);
                        rep[0][1] = say * (A.gneg * ra);
                        rep[1][0] = sbx * (B.gneg * rb);
                        rep[1][1] = sby * (B.gneg * rb);
                        // adhesion: weighted lattice vectors towards solid neighbours
                        double ssx = 0.0, ssy = 0.0;
#pragma unroll
                        for (int i = 1; i < Q; ++i) {
                            if (!((nbc >> i) & 1u)) {
                                ssx += wint(i) * cx(i);
                                ssy += wint(i) * cy(i);
                            }
                        }
                        const double fs[2][2] = {{A.gsneg * ssx, A.gsneg * ssy},
                                                 {B.gsneg * ssx, B.gsneg * ssy}};
                        momentum(pc[0], jm[0][0], jm[0][1]);
                        momentum(pc[1], jm[1][0], jm[1][1]);
                        const double den = ra * A.omega + rb * B.omega;
                        const double ucx = (jm[0][0] * A.omega + jm[1][0] * B.omega) / den;
                        const double ucy = (jm[0][1] * A.omega + jm[1][1] * B.omega) / den;
#pragma unroll
                        for (int c = 0; c < 2; ++c) {
                            const auto& C = a.c[c];
                            const double r = rc[c];
                            if (!is_dry(r)) {
                                ue[c][0] = ucx + rep[c][0] * C.inv_omega / r + fs[c][0] * C.inv_omega;
                                ue[c][1] = ucy + rep[c][1] * C.inv_omega / r + fs[c][1] * C.inv_omega;
                                if (!REPLAY && fabs(r) < kUeqWatch) {
                                    C.ueqx[s_idx] = ue[c][0];
                                    C.ueqy[s_idx] = ue[c][1];
                                }
                            } else {
                                ue[c][0] = C.ueqx[s_idx];
                                ue[c][1] = C.ueqy[s_idx];
                            }
                        }
                    } else if (REPLAY) {
                        ue[0][0] = A.ueqx[s_idx];
                        ue[0][1] = A.ueqy[s_idx];
                        ue[1][0] = B.ueqx[s_idx];
                        ue[1][1] = B.ueqy[s_idx];
                    }
                    if (fl || REPLAY) {
#pragma unroll
                        for (int c = 0; c < 2; ++c) feq_ref(rc[c], ue[c][0], ue[c][1], e[c]);
                    }
                    if (REPLAY) {
#pragma unroll
                        for (int c = 0; c < 2; ++c) {
                            const auto& C = a.c[c];
                            C.rho[s_idx] = rc[c];
                            C.jx[s_idx] = jm[c][0];
                            C.jy[s_idx] = jm[c][1];
                            C.frx[s_idx] = rep[c][0];
                            C.fry[s_idx] = rep[c][1];
                            if (fl) {
                                C.ueqx[s_idx] = ue[c][0];
                                C.ueqy[s_idx] = ue[c][1];
                            }
#pragma unroll
                            for (int i = 0; i < Q; ++i) C.feq[i * plane + s_idx] = e[c][i];
                        }
                    } else if (fl) {
#pragma unroll
                        for (int c = 0; c < 2; ++c) {
                            const auto& C = a.c[c];
                            S* d = dst[c];
#pragma unroll
                            for (int i = 0; i < Q; ++i) {
                                double v = C.omc * pc[c][i] + C.omega * e[c][i];
                                if (C.has_force) v += C.force[i];
                                ST::st(d + i * plane + s_idx, v);
                            }
                        }
                    }
                } else {
                    // deviation forms, FP32 arithmetic
                    if (!REPLAY && fl) {
                        const auto& A = a.c[0];
                        const auto& B = a.c[1];
                        float sd[2][2] = {{0.f, 0.f}, {0.f, 0.f}};   // sums over fluid nbrs of w c drho_other
                        float sfx = 0.f, sfy = 0.f;                  // sums over fluid nbrs of w c
#pragma unroll
                        for (int i = 1; i < Q; ++i) {
                            if ((nbc >> i) & 1u) {
                                const int sl = cy(i) == 1 ? sN : (cy(i) == -1 ? sS : sC);
                                const float da = lds[sl][0][cl + cx(i)];
                                const float db = lds[sl][1][cl + cx(i)];
                                const float wx = (float)(wint(i) * cx(i)), wy = (float)(wint(i) * cy(i));
                                sd[0][0] += wx * db;
                                sd[0][1] += wy * db;
                                sd[1][0] += wx * da;
                                sd[1][1] += wy * da;
                                sfx += wx;
                                sfy += wy;
                            }
                        }
                        const float da0 = rc[0], db0 = rc[1];
                        const float ra = A.rho0 + da0, rb = B.rho0 + db0;
                        // psi_other summed over fluid neighbours = sum w c drho + rho0_other * sum w c
                        const float fa_x = A.gneg * ra * (sd[0][0] + B.rho0 * sfx);
                        const float fa_y = A.gneg * ra * (sd[0][1] + B.rho0 * sfy);
                        const float fb_x = B.gneg * rb * (sd[1][0] + A.rho0 * sfx);
                        const float fb_y = B.gneg * rb * (sd[1][1] + A.rho0 * sfy);
                        // adhesion: sum over solid nbrs of w c = -(sum over fluid nbrs)
                        const float fs_ax = A.gsneg * -sfx, fs_ay = A.gsneg * -sfy;
                        const float fs_bx = B.gsneg * -sfx, fs_by = B.gsneg * -sfy;
                        float jax, jay, jbx, jby;
                        momentum(pc[0], jax, jay);
                        momentum(pc[1], jbx, jby);
                        const float den = ra * A.omega + rb * B.omega;
                        const float ucx = (jax * A.omega + jbx * B.omega) / den;
                        const float ucy = (jay * A.omega + jby * B.omega) / den;
                        float uea[2], ueb[2];
                        if (!is_dry(ra)) {
                            uea[0] = ucx + A.inv_omega * (fa_x / ra + fs_ax);
                            uea[1] = ucy + A.inv_omega * (fa_y / ra + fs_ay);
                            if (fabsf(ra) < (float)kUeqWatch) {
                                A.ueqx[s_idx] = uea[0];
                                A.ueqy[s_idx] = uea[1];
                            }
                        } else {
                            uea[0] = (float)A.ueqx[s_idx];
                            uea[1] = (float)A.ueqy[s_idx];
                        }
                        if (!is_dry(rb)) {
                            ueb[0] = ucx + B.inv_omega * (fb_x / rb + fs_bx);
                            ueb[1] = ucy + B.inv_omega * (fb_y / rb + fs_by);
                            if (fabsf(rb) < (float)kUeqWatch) {
                                B.ueqx[s_idx] = ueb[0];
                                B.ueqy[s_idx] = ueb[1];
                            }
                        } else {
                            ueb[0] = (float)B.ueqx[s_idx];
                            ueb[1] = (float)B.ueqy[s_idx];
                        }
                        float ea[Q], eb[Q];
                        feq_dev(ra, da0, uea[0], uea[1], ea);
                        feq_dev(rb, db0, ueb[0], ueb[1], eb);
                        S* d0 = dst[0];
                        S* d1 = dst[1];
#pragma unroll
                        for (int i = 0; i < Q; ++i) {
                            float va = A.omc * pc[0][i] + A.omega * ea[i];
                            float vb = B.omc * pc[1][i] + B.omega * eb[i];
                            if (A.has_force) va += A.force[i];
                            if (B.has_force) vb += B.force[i];
                            ST::st(d0 + i * plane + s_idx, va);
                            ST::st(d1 + i * plane + s_idx, vb);
                        }
                    }
                }
            }
        }
        // rotate registers
        if (xin) {
            nbc = nbn;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                rc[c] = rn[c];
#pragma unroll
                for (int i = 0; i < Q; ++i) pc[c][i] = pn[c][i];
            }
        }
    }
}

}  // namespace scl
