// serve.h -- device side of the serving layer: the persistent multi-job step
// kernel over a tile work-list, the per-job mass reductions of the watchdog,
// and the per-job set-up and delivery kernels.
//
// Work-list.  A tile is (job, strip of 64*NW columns, band of rows [y0, y1)).
// Every step of a round is one launch of k_serve with a fixed grid (the
// workgroups that fit on the device at once); each workgroup takes tiles in
// order from a ticket counter until the list is exhausted, so small and
// large jobs share the device without per-job launches.  A tile is computed
// by step_tile, the body of the single-lattice fused kernel, so a job's
// populations are the same bits whichever tiles, workgroups or other jobs
// share its steps.
//
// The ticket counter resets itself: after its last fetch each workgroup
// increments a second counter, and the workgroup that completes it zeroes
// both, so consecutive launches (and replays of a captured graph) start from
// zero without a memset.
//
// Two or three steps per launch (deviation forms, jobs of at least 8 columns
// and 2K + 2 rows with K or more steps in the round): k_serve2x / k_serve3x
// take tiles of a second / third list (strips of the K-step tile's output
// columns) and compute each with stepk_tile, the body of the single-lattice
// multi-step kernels; launch q of a round advances steps Kq .. Kq + K - 1 of
// such a job.  The steps past the last whole group of K (one for an odd
// two-step job, one or two for a three-step job) go through k_serve.  A
// launch reads one buffer and writes the other, so after q launches the
// job's populations are in buffer cur ^ (q & 1).  Rounds in which no such
// job has a body force launch the force-free variants.
#pragma once
#include "../host/serve_ic.h"
#include "convert.h"
#include "step.h"
#include "stepk.h"

namespace scl {

// One resident job as the kernels see it.  a.c[c].src / a.c[c].dst are
// buffers 0 / 1 of component c; `cur` names the buffer that holds the job's
// post-collision populations at the start of the round, so step s reads
// buffer cur ^ (s & 1).
template <class T>
struct JobArgs {
    StepArgs<T> a;
    int cur;
    int steps;                  // steps of this job in the current round
    int x2;                     // 1: steps 2q, 2q+1 go through k_serve2x (q < steps / 2)
    int x3;                     // 1: steps 3q .. 3q+2 go through k_serve3x (q < steps / 3)
    int fin;                    // buffer holding the populations at the end of the round
    int item_begin, item_end;   // its tiles in the work-list
    // deviation forms: rotating equilibrium-velocity buffers [k][component]
    // [x/y]; step s reads buffer (ucur + s) % kUeqRing and writes the next
    double* ueq[kUeqRing][2][2];
    int ucur;
    UeqRing ring;               // the same buffers as a ring (cur = ucur)
};

struct ServeHdr {
    const void* jobs;           // JobArgs<T>[n_jobs]
    const uint64_t* items;      // work-list: every job's tiles (the mass reduction's)
    double* partial;            // 2 per tile (mass per component)
    double* mass;               // 2 per job
    int n_items, n_jobs;
    int bw;                     // tile width in columns
    const uint64_t* items2;     // two-step tiles (strips of x2_out columns)
    int n_items2;
    const uint64_t* items3;     // three-step tiles (strips of xk_out<ST, 3, x3_threads> columns)
    int n_items3;
    unsigned* flags;            // per job: the zero-density rule's flags of both components
    // per step s of the round, three ints P, B, E: k_serve's launch s takes
    // items [0, P) (the single-step jobs with more than s steps) and [B, E)
    // (the multi-step jobs that take step s singly)
    const int* sched;
};

// Tile encoding, 16 bits each from the low end: job, strip, y0, y1.
__host__ __device__ inline uint64_t pack_item(int job, int bx, int y0, int y1) {
    return (uint64_t)(unsigned)job | ((uint64_t)(unsigned)bx << 16) | ((uint64_t)(unsigned)y0 << 32) |
           ((uint64_t)(unsigned)y1 << 48);
}

__device__ inline void unpack_item(uint64_t it, int& job, int& bx, int& y0, int& y1) {
    job = (int)(it & 0xFFFFu);
    bx = (int)((it >> 16) & 0xFFFFu);
    y0 = (int)((it >> 32) & 0xFFFFu);
    y1 = (int)(it >> 48);
}

template <class ST, int NW, int LA>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(step_waves_per_eu<ST, false, LA>()))) void
k_serve(const ServeHdr* __restrict__ hdr, int s, unsigned* __restrict__ ctr) {
    using T = typename ST::T;
    __shared__ DensRing<2, T, NW> lds;
    __shared__ int ticket;
    const SCL_AS_CONST int* sched = K(hdr->sched);
    const int P = sched[3 * s], B = sched[3 * s + 1], n = P + (sched[3 * s + 2] - B);
    if (n == 0) return;   // nothing at step s: the counters stay zero untouched
    const SCL_AS_CONST JobArgs<T>* jobs = K((const JobArgs<T>*)hdr->jobs);
    const SCL_AS_CONST uint64_t* items = K(hdr->items);
    for (;;) {
        __syncthreads();   // the previous tile is done with `ticket` and the ring
        if (threadIdx.x == 0)
            ticket = (int)__hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(ticket);
        if (t >= n) break;
        int j, bx, y0, y1;
        unpack_item(items[t < P ? t : B + (t - P)], j, bx, y0, y1);
        j = __builtin_amdgcn_readfirstlane(j);
        bx = __builtin_amdgcn_readfirstlane(bx);
        y0 = __builtin_amdgcn_readfirstlane(y0);
        y1 = __builtin_amdgcn_readfirstlane(y1);
        const SCL_AS_CONST JobArgs<T>& J = jobs[j];
        if (s >= J.steps) continue;
        if (J.x2 && s != J.steps - 1) continue;   // the pairs are k_serve2x's
        if (J.x3 && s < J.steps / 3 * 3) continue;   // the triples are k_serve3x's
        StepArgs<T> a = J.a;
#pragma unroll
        for (int c = 0; c < 2; ++c) a.c[c].step = J.a.c[c].step + (unsigned)s;
        if constexpr (ST::kDev) {
            const int ki = (J.ucur + s) % kUeqRing, ko = (J.ucur + s + 1) % kUeqRing;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                a.c[c].ueqx_in = J.ueq[ki][c][0];
                a.c[c].ueqy_in = J.ueq[ki][c][1];
                a.c[c].ueqx = J.ueq[ko][c][0];
                a.c[c].ueqy = J.ueq[ko][c][1];
            }
        }
        Bufs b;
        // buffer read at step s: one swap per launch (after s / 2 pairs for a
        // two-step job, after steps / 3 triples and s - 3 (steps / 3) single
        // steps for a three-step one)
        const int sw = (J.cur ^ (J.x2 ? (s >> 1) : (J.x3 ? s - 2 * (J.steps / 3) : s))) & 1;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            b.src[c] = sw ? a.c[c].dst : a.c[c].src;
            b.dst[c] = sw ? const_cast<void*>(a.c[c].src) : a.c[c].dst;
        }
        step_tile<2, ST, NW, true, false, uint32_t, LA, false>(a, b, bx, y0, y1, lds);
    }
    if (threadIdx.x == 0) {
        // every fetch of this workgroup precedes this release
        const unsigned d = __hip_atomic_fetch_add(ctr + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (d == gridDim.x - 1) {
            __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(ctr + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Launch q of a round: steps 2q and 2q+1 of every two-step job that has them,
// one tile of the second list per ticket (the general variant, since the
// jobs carry masks).  FORCE = false: no two-step job of the round has a body
// force.
template <class ST, bool FORCE>
__global__ __launch_bounds__(kX2Threads) void k_serve2x(const ServeHdr* __restrict__ hdr, int q,
                                                        unsigned* __restrict__ ctr) {
    __shared__ XkLds<ST, false, 2, kX2Threads> lds;
    __shared__ int ticket;
    const int n = hdr->n_items2;
    const SCL_AS_CONST JobArgs<float>* jobs = K((const JobArgs<float>*)hdr->jobs);
    const SCL_AS_CONST uint64_t* items = K(hdr->items2);
    for (;;) {
        __syncthreads();   // the previous tile is done with `ticket` and the LDS
        if (threadIdx.x == 0)
            ticket = (int)__hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(ticket);
        if (t >= n) break;
        int j, bx, y0, y1;
        unpack_item(items[t], j, bx, y0, y1);
        j = __builtin_amdgcn_readfirstlane(j);
        bx = __builtin_amdgcn_readfirstlane(bx);
        y0 = __builtin_amdgcn_readfirstlane(y0);
        y1 = __builtin_amdgcn_readfirstlane(y1);
        const SCL_AS_CONST JobArgs<float>& J = jobs[j];
        if (2 * q + 1 >= J.steps) continue;
        XkArgs<2> A;
        A.a = J.a;
        A.u = J.ring;
        const int rb = (J.cur ^ q) & 1;
        const int cur = (J.ucur + 2 * q) % J.ring.nsets;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            A.a.c[c].src = rb ? J.a.c[c].dst : J.a.c[c].src;
            A.dst[c] = rb ? const_cast<void*>(J.a.c[c].src) : J.a.c[c].dst;
            A.u.cur[c] = cur;
        }
        stepk_tile<ST, false, 2, kX2Threads, FORCE>(A, bx, y0, y1, lds);
    }
    if (threadIdx.x == 0) {
        const unsigned d = __hip_atomic_fetch_add(ctr + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (d == gridDim.x - 1) {
            __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(ctr + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Launch q of a round: steps 3q .. 3q+2 of every three-step job that has
// them (the tile with solids, strips of x3_threads columns).  FORCE = false:
// no three-step job of the round has a body force.  Two waves per SIMD are
// asked for: left free, the compiler gave the FP16 variant with a force 256
// vector and 2 accumulation registers, one wave per SIMD.
template <class ST, bool FORCE>
__global__ __launch_bounds__((x3_threads<ST, false>())) __attribute__((amdgpu_waves_per_eu(2))) void k_serve3x(const ServeHdr* __restrict__ hdr, int q,
                                                                     unsigned* __restrict__ ctr) {
    constexpr int T3 = x3_threads<ST, false>();
    __shared__ XkLds<ST, false, 3, T3> lds;
    __shared__ int ticket;
    const int n = hdr->n_items3;
    const SCL_AS_CONST JobArgs<float>* jobs = K((const JobArgs<float>*)hdr->jobs);
    const SCL_AS_CONST uint64_t* items = K(hdr->items3);
    for (;;) {
        __syncthreads();   // the previous tile is done with `ticket` and the LDS
        if (threadIdx.x == 0)
            ticket = (int)__hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(ticket);
        if (t >= n) break;
        int j, bx, y0, y1;
        unpack_item(items[t], j, bx, y0, y1);
        j = __builtin_amdgcn_readfirstlane(j);
        bx = __builtin_amdgcn_readfirstlane(bx);
        y0 = __builtin_amdgcn_readfirstlane(y0);
        y1 = __builtin_amdgcn_readfirstlane(y1);
        const SCL_AS_CONST JobArgs<float>& J = jobs[j];
        if (3 * q + 2 >= J.steps) continue;
        XkArgs<3> A;
        A.a = J.a;
        A.u = J.ring;
        const int rb = (J.cur ^ q) & 1;
        const int cur = (J.ucur + 3 * q) % J.ring.nsets;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            A.a.c[c].src = rb ? J.a.c[c].dst : J.a.c[c].src;
            A.dst[c] = rb ? const_cast<void*>(J.a.c[c].src) : J.a.c[c].dst;
            A.u.cur[c] = cur;
        }
        stepk_tile<ST, false, 3, T3, FORCE>(A, bx, y0, y1, lds);
    }
    if (threadIdx.x == 0) {
        const unsigned d = __hip_atomic_fetch_add(ctr + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (d == gridDim.x - 1) {
            __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(ctr + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Sum of two doubles per thread over a 256-thread workgroup, fixed order.
__device__ inline void block_sum2(double& a, double& b, double (&red)[2][256]) {
    const int t = threadIdx.x;
    red[0][t] = a;
    red[1][t] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            red[0][t] += red[0][t + w];
            red[1][t] += red[1][t + w];
        }
        __syncthreads();
    }
    a = red[0][0];
    b = red[1][0];
}

// Watchdog mass: FP64 sum over the fluid sites of each tile of the
// post-collision populations the round ended with (collision and the
// force vectors conserve each site's mass and the gather permutes slots, so
// this equals the mass of the streamed state up to rounding).
template <class ST>
__global__ __launch_bounds__(256) void k_serve_mass(const ServeHdr* __restrict__ hdr) {
    using T = typename ST::T;
    __shared__ double red[2][256];
    const int n = hdr->n_items, bw = hdr->bw;
    const JobArgs<T>* __restrict__ jobs = (const JobArgs<T>*)hdr->jobs;
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        int j, bx, y0, y1;
        unpack_item(hdr->items[t], j, bx, y0, y1);
        const JobArgs<T>& J = jobs[j];
        const int fin = J.fin;
        const typename ST::S* P[2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
            P[c] = (const typename ST::S*)(fin ? (const void*)J.a.c[c].dst : J.a.c[c].src);
        const StepArgs<T>& a = J.a;
        const int x0 = bx * bw, w = min(bw, a.nx - x0);
        double m[2] = {0.0, 0.0};
        for (int k = threadIdx.x; k < w * (y1 - y0); k += 256) {
            const int x = x0 + k % w, y = y0 + k / w;
            if (!fluid_bit(a.mask, a.wpr, y, x)) continue;
            const long long s_idx = (long long)y * a.pitch + x;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                double r = ST::kDev ? (double)a.c[c].rho0 : 0.0;
#pragma unroll
                for (int i = 0; i < Q; ++i) r += (double)ST::ld(P[c] + i * a.plane + s_idx);
                m[c] += r;
            }
        }
        block_sum2(m[0], m[1], red);
        if (threadIdx.x == 0) {
            hdr->partial[2 * t] = m[0];
            hdr->partial[2 * t + 1] = m[1];
        }
        __syncthreads();
    }
}

// Per-job sums of the tile partials, in tile order.
template <class T>
__global__ __launch_bounds__(256) void k_serve_mass_final(const ServeHdr* __restrict__ hdr) {
    __shared__ double red[2][256];
    const int j = blockIdx.x;
    if (j >= hdr->n_jobs) return;
    const JobArgs<T>& J = ((const JobArgs<T>*)hdr->jobs)[j];
    double m0 = 0.0, m1 = 0.0;
    for (int t = J.item_begin + threadIdx.x; t < J.item_end; t += 256) {
        m0 += hdr->partial[2 * t];
        m1 += hdr->partial[2 * t + 1];
    }
    block_sum2(m0, m1, red);
    if (threadIdx.x == 0) {
        const long long pl = J.a.plane;   // FP64 jobs keep the zero-density rule's records
        hdr->flags[j] = J.a.c[0].stamp ? G(J.a.c[0].stamp)[pl] | G(J.a.c[1].stamp)[pl] : 0u;
        hdr->mass[2 * j] = m0;
        hdr->mass[2 * j + 1] = m1;
    }
}

// Generated initial state S_0 of one component in the storage form:
// f_i = rho/9 on fluid sites (the reference initialisers, src/lbm.cpp:33-46,
// src/fluid.cpp:96-107), 0 on solid sites; deviation forms store
// f_i - w_i*rho0 converted as k_from_f64 does.
template <class ST>
__global__ void k_init_ic(Grid2 g, serve::IcParams p, int c, double rho0, typename ST::S* __restrict__ dst) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    const bool fl = fluid_bit(g.mask, g.wpr, y, x);
    const double fi = fl ? serve::ic_density_at(p, c, x, y) / 9 : 0.0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        if (ST::kDev) ST::st(dst + i * g.plane + s, (typename ST::T)(fi - weq(i) * rho0));
        else ST::st(dst + i * g.plane + s, (typename ST::T)fi);
    }
}

// Inverse of the gather: post-collision populations P whose gather is the
// streamed state S.  Slot i of fluid site z is read by site z + c_i when that
// site is fluid, and by z itself as direction opp(i) otherwise, so
//   P_i(z) = S_i(z + c_i)   if z + c_i is fluid,
//          = S_opp(i)(z)    otherwise.
// The gathering kernels read the bounce-back value of a fluid site x from
// slot j of its solid source s = x - c_j (the ghost slot the stepping kernel
// writes), so the ghost slots of solid sites are filled too:
//   P_j(s) = S_j(s + c_j)   for solid s and fluid s + c_j.
// Pure copies: Gather(Scatter(S)) == S bit for bit on fluid sites.
template <class ST>
__global__ void k_scatter(Grid2 g, const typename ST::S* __restrict__ src, typename ST::S* __restrict__ dst) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    const unsigned nb = nbits_at(g, x, y);
    const int xm = wrap1(x - 1, g.nx), xp = wrap1(x + 1, g.nx);
    const int ym = wrap1(y - 1, g.ny), yp = wrap1(y + 1, g.ny);
    const bool fl = nb & 1u;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int sx = cx(i) == 1 ? xp : (cx(i) == -1 ? xm : x);
        const int sy = cy(i) == 1 ? yp : (cy(i) == -1 ? ym : y);
        const long long nbr = (long long)i * g.plane + (long long)sy * g.pitch + sx;   // slot i of z + c_i
        if (fl) {
            dst[i * g.plane + s] = ((nb >> i) & 1u) ? src[nbr] : src[(long long)opp(i) * g.plane + s];
        } else if (i > 0 && ((nb >> i) & 1u)) {
            dst[i * g.plane + s] = src[nbr];
        }
    }
}

// Density of the streamed state Gather(P) in FP64, reference order
// (src/fluid.cpp:177-182) over the FP64 values of the populations; 0 on
// solid sites.
template <class ST>
__global__ void k_job_density(Grid2 g, const typename ST::S* __restrict__ P, double rho0, double* __restrict__ rho) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    const unsigned nb = nbits_at(g, x, y);
    double r = 0.0;
    if (nb & 1u) {
        typename ST::T p[Q];
        load_site<ST, true>(P, g.plane, g.pitch, x, wrap1(x - 1, g.nx), wrap1(x + 1, g.nx), y, wrap1(y - 1, g.ny),
                            wrap1(y + 1, g.ny), nb, p);
#pragma unroll
        for (int i = 0; i < Q; ++i) r += ST::kDev ? (double)p[i] + weq(i) * rho0 : (double)p[i];
    }
    rho[s] = r;
}

}  // namespace scl
