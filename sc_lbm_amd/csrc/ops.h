// ops.h -- per-operator kernels (reference semantics, one pass each) and the
// helper kernels around the fused step: gather (stream only), storage
// conversions and FP64 moment reductions.
//
// The per-operator kernels back the reference's fine-grained API (the
// individual LBM / Fluid methods) when a caller does not issue the canonical
// step sequence.  They work on the reference's streamed state S in FP64 and
// reproduce each method's arithmetic as written (-ffp-contract=off).
#pragma once
#include "convert.h"
#include "lattice.h"
#include "step.h"

namespace scl {

// Fluid::compute_density, src/fluid.cpp:175-183 (every site).
__global__ void k_density(Grid2 g, const double* __restrict__ f, double* __restrict__ rho) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < Q; ++i) r += f[i * g.plane + s];
    rho[s] = r;
}

// Fluid::compute_velocities, src/fluid.cpp:186-200.
__global__ void k_velocities(Grid2 g, const double* __restrict__ f, const double* __restrict__ rho,
                             double* __restrict__ ux, double* __restrict__ uy) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    double vx = 0.0, vy = 0.0;
    if (fluid_bit(g.mask, g.wpr, y, x)) {
        double p[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) p[i] = f[i * g.plane + s];
        momentum(p, vx, vy);
        vx /= rho[s];
        vy /= rho[s];
    }
    ux[s] = vx;
    uy[s] = vy;
}

// feq on every site from rho and a velocity (src/fluid.cpp:217-242 / 254-276).
__global__ void k_feq(Grid2 g, const double* __restrict__ rho, const double* __restrict__ ux,
                      const double* __restrict__ uy, double* __restrict__ feq) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    double e[Q];
    feq_ref(rho[s], ux[s], uy[s], e);
#pragma unroll
    for (int i = 0; i < Q; ++i) feq[i * g.plane + s] = e[i];
}

// LBM::compute_solid_surface_force + Fluid::add_surface_forces
// (src/lbm.cpp:102-141, src/fluid.cpp:154-160) for one component.
__global__ void k_solid_force(Grid2 g, double gsneg, double* __restrict__ fsx, double* __restrict__ fsy) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    double sx = 0.0, sy = 0.0;
    const unsigned nb = nbits_at(g, x, y);
    if (nb & 1u) {
#pragma unroll
        for (int i = 1; i < Q; ++i) {
            if (!((nb >> i) & 1u)) {
                sx += wint(i) * cx(i);
                sy += wint(i) * cy(i);
            }
        }
    }
    fsx[s] = gsneg * sx;
    fsy[s] = gsneg * sy;
}

// LBM::compute_fluid_repulsive_interactions, src/lbm.cpp:144-206.
__global__ void k_repulsion(Grid2 g, const double* __restrict__ ra, const double* __restrict__ rb,
                            double gnega, double gnegb, double* __restrict__ fax, double* __restrict__ fay,
                            double* __restrict__ fbx, double* __restrict__ fby) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
    const unsigned nb = nbits_at(g, x, y);
    if (nb & 1u) {
#pragma unroll
        for (int i = 1; i < Q; ++i) {
            if ((nb >> i) & 1u) {
                const int xn = wrap1(x + cx(i), g.nx), yn = wrap1(y + cy(i), g.ny);
                const int t = yn * g.pitch + xn;
                const double pa = ra[t], pb = rb[t];
                if (cx(i) != 0) {
                    ax += (wint(i) * cx(i)) * pb;
                    bx += (wint(i) * cx(i)) * pa;
                }
                if (cy(i) != 0) {
                    ay += (wint(i) * cy(i)) * pb;
                    by += (wint(i) * cy(i)) * pa;
                }
            }
        }
        ax *= gnega * ra[s];
        ay *= gnega * ra[s];
        bx *= gnegb * rb[s];
        by *= gnegb * rb[s];
    }
    fax[s] = ax;
    fay[s] = ay;
    fbx[s] = bx;
    fby[s] = by;
}

struct EqVelComp {
    const double* f;
    const double* rho;
    double omega, inv_omega;
    const double *frx, *fry, *fsx, *fsy;
    double *jx, *jy, *ueqx, *ueqy;
};

// LBM::compute_equilibrium_velocities, src/lbm.cpp:209-272.
__global__ void k_eq_velocities(Grid2 g, EqVelComp A, EqVelComp B) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    double j[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    if (fluid_bit(g.mask, g.wpr, y, x)) {
        double p[Q];
        const EqVelComp* C[2] = {&A, &B};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
#pragma unroll
            for (int i = 0; i < Q; ++i) p[i] = C[c]->f[i * g.plane + s];
            momentum(p, j[c][0], j[c][1]);
        }
        const double ra = A.rho[s], rb = B.rho[s];
        const double ucx = (j[0][0] * A.omega + j[1][0] * B.omega) / (ra * A.omega + rb * B.omega);
        const double ucy = (j[0][1] * A.omega + j[1][1] * B.omega) / (ra * A.omega + rb * B.omega);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const EqVelComp& Cc = *C[c];
            const double r = Cc.rho[s];
            if (!is_dry(r)) {
                Cc.ueqx[s] = ucx + Cc.frx[s] * Cc.inv_omega / r + Cc.fsx[s] * Cc.inv_omega;
                Cc.ueqy[s] = ucy + Cc.fry[s] * Cc.inv_omega / r + Cc.fsy[s] * Cc.inv_omega;
            }
        }
    }
    A.jx[s] = j[0][0];
    A.jy[s] = j[0][1];
    B.jx[s] = j[1][0];
    B.jy[s] = j[1][1];
}

// BGK relaxation, src/lbm.cpp:283-287 / 306-310, every site.
__global__ void k_bgk(Grid2 g, double* __restrict__ f, const double* __restrict__ feq, double omc, double omega) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const long long k = i * g.plane + s;
        f[k] = omc * f[k] + omega * feq[k];
    }
}

struct Force9 {
    double v[9];
};

// LBM::add_volume_force, src/lbm.cpp:315-325 (fluid sites).
__global__ void k_volume_force(Grid2 g, double* __restrict__ f, Force9 F) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    if (!fluid_bit(g.mask, g.wpr, y, x)) return;
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i * g.plane + s] += F.v[i];
}

// FP64 mass and momentum of FP64 populations (pitched): one partial triple
// per row, then a single-workgroup pass sums the rows in order.
__global__ void k_moments_rows(Grid2 g, const double* __restrict__ f, double* __restrict__ part) {
    const int y = blockIdx.x;
    double m = 0.0, jx = 0.0, jy = 0.0;
    for (int x = threadIdx.x; x < g.nx; x += blockDim.x) {
        const int s = y * g.pitch + x;
        double p[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) p[i] = f[i * g.plane + s];
        double mx, my;
        momentum(p, mx, my);
        m += dens(p);
        jx += mx;
        jy += my;
    }
    __shared__ double red[3][256];
    red[0][threadIdx.x] = m;
    red[1][threadIdx.x] = jx;
    red[2][threadIdx.x] = jy;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) part[3 * y + k] = red[k][0];
    }
}

// The same per-row partial triples as k_to_f64 followed by k_moments_rows
// (the same per-site FP64 values, summed in the same order), read from the
// populations as stored, gathered from the post-collision phase when GATHER:
// no FP64 staging copy (scl_moments_noalloc).
template <class ST, bool GATHER>
__global__ void k_moments_direct(Grid2 g, const typename ST::S* __restrict__ src, double rho0,
                                 double* __restrict__ part) {
    const int y = blockIdx.x;
    double m = 0.0, jx = 0.0, jy = 0.0;
    for (int x = threadIdx.x; x < g.nx; x += blockDim.x) {
        const unsigned nb = nbits_at(g, x, y);
        const bool fl = nb & 1u;
        typename ST::T q[Q];
        if (!GATHER || fl)
            load_site<ST, GATHER>(src, g.plane, g.pitch, x, wrap1(x - 1, g.nx), wrap1(x + 1, g.nx), y,
                                  wrap1(y - 1, g.ny), wrap1(y + 1, g.ny), nb, q);
        double p[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (GATHER && !fl) p[i] = 0.0;
            else if (ST::kDev) p[i] = (double)q[i] + weq(i) * rho0;
            else p[i] = (double)q[i];
        }
        double mx, my;
        momentum(p, mx, my);
        m += dens(p);
        jx += mx;
        jy += my;
    }
    __shared__ double red[3][256];
    red[0][threadIdx.x] = m;
    red[1][threadIdx.x] = jx;
    red[2][threadIdx.x] = jy;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) part[3 * y + k] = red[k][0];
    }
}

__global__ void k_moments_final(int rows, const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double red[3][256];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] += part[3 * r + k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k] = red[k][0];
    }
}

// Droplet observables (the reference evaluates them in Python from the
// density files its drivers write).  Per row y of the FP64 populations of
// the droplet (fd) and bulk (fb) components, the number of sites and the
// least and greatest x of two site sets:
//   out[6y + 0..2]: the Laplace-law set, rho_d - rho_b < 0, rho_b <= rho_bulk,
//     rho_b > rho_disp (laplace_law/compute_surface_tension.py:66);
//   out[6y + 3..5]: the contact-angle interface, rho_b - rho_d <= 0,
//     rho_d <= rho_bulk, rho_b > rho_disp (contact_angles/
//     compute_contact_angles.py:47).
// Densities in the reference's summation order (dens).  Empty sets leave
// min x = nx, max x = -1.
__global__ void k_obs_rows(Grid2 g, const double* __restrict__ fd, const double* __restrict__ fb, double rho_bulk,
                           double rho_disp, int* __restrict__ out) {
    const int y = blockIdx.x;
    int n[2] = {0, 0}, lo[2] = {g.nx, g.nx}, hi[2] = {-1, -1};
    for (int x = threadIdx.x; x < g.nx; x += blockDim.x) {
        const int s = y * g.pitch + x;
        double pd[Q], pb[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            pd[i] = fd[i * g.plane + s];
            pb[i] = fb[i * g.plane + s];
        }
        const double rd = dens(pd), rb = dens(pb);
        const bool in[2] = {rd - rb < 0.0 && rb <= rho_bulk && rb > rho_disp,
                            rb - rd <= 0.0 && rd <= rho_bulk && rb > rho_disp};
#pragma unroll
        for (int k = 0; k < 2; ++k)
            if (in[k]) {
                ++n[k];
                lo[k] = min(lo[k], x);
                hi[k] = max(hi[k], x);
            }
    }
    __shared__ int red[6][256];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        red[3 * k][threadIdx.x] = n[k];
        red[3 * k + 1][threadIdx.x] = lo[k];
        red[3 * k + 2][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                red[3 * k][threadIdx.x] += red[3 * k][threadIdx.x + w];
                red[3 * k + 1][threadIdx.x] = min(red[3 * k + 1][threadIdx.x], red[3 * k + 1][threadIdx.x + w]);
                red[3 * k + 2][threadIdx.x] = max(red[3 * k + 2][threadIdx.x], red[3 * k + 2][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[6 * y + threadIdx.x] = red[threadIdx.x][0];
}

// Densities of both components at sites s0 and s1: out = {rho_d(s0),
// rho_b(s0), rho_d(s1), rho_b(s1)}.  One thread.
__global__ void k_obs_points(Grid2 g, const double* __restrict__ fd, const double* __restrict__ fb, int s0, int s1,
                             double* __restrict__ out) {
    const int ss[2] = {s0, s1};
    for (int k = 0; k < 2; ++k) {
        double pd[Q], pb[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            pd[i] = fd[i * g.plane + ss[k]];
            pb[i] = fb[i * g.plane + ss[k]];
        }
        out[2 * k] = dens(pd);
        out[2 * k + 1] = dens(pb);
    }
}

// Populations of a fluid at rest with density rho everywhere (f_i = w_i*rho),
// in the storage form; used for the autotuner's scratch fields.
template <class ST>
__global__ void k_fill_rest(Grid2 g, typename ST::S* __restrict__ dst, double rho, double rho0) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const double v = weq(i) * rho - (ST::kDev ? weq(i) * rho0 : 0.0);
        ST::st(dst + i * g.plane + s, (typename ST::T)v);
    }
}

// Plain copy used to calibrate the HBM byte counters and to measure the
// practical read+write bandwidth: each lane moves four 16-byte vectors that
// are all in flight together; the grid covers the buffer once.
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_probe_copy(const f32x4* __restrict__ src, f32x4* __restrict__ dst,
                                                    size_t n) {
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    f32x4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const size_t i = base + 256 * k;
        if (i < n) v[k] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const size_t i = base + 256 * k;
        if (i < n) __builtin_nontemporal_store(v[k], dst + i);
    }
}

}  // namespace scl
