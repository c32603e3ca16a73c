// convert.h -- site-parallel helpers shared by the library's translation
// units: the lattice descriptor of the one-thread-per-site kernels, the
// stream-only gather and the conversions between FP64 populations and the
// storage forms.  Templates and inline functions only, so any number of
// translation units can include it.
#pragma once
#include "lattice.h"
#include "step.h"

namespace scl {

struct Grid2 {
    int nx, ny, pitch, wpr;
    long long plane;
    const uint64_t* mask;
};

__device__ inline bool site_of(const Grid2& g, int& x, int& y, int& s) {
    x = blockIdx.x * blockDim.x + threadIdx.x;
    y = blockIdx.y;
    s = y * g.pitch + x;
    return x < g.nx;
}

__device__ inline unsigned nbits_at(const Grid2& g, int x, int y) {
    const int xm = wrap1(x - 1, g.nx), xp = wrap1(x + 1, g.nx);
    const int ym = wrap1(y - 1, g.ny), yp = wrap1(y + 1, g.ny);
    return neigh_bits(triple(g.mask, g.wpr, ym, xm, x, xp), triple(g.mask, g.wpr, y, xm, x, xp),
                      triple(g.mask, g.wpr, yp, xm, x, xp));
}

// Stream with link bounce-back as a pull (src/lbm.cpp:343-435 seen from the
// receiving site); solid sites become 0.  Works for any storage form (the
// deviation forms store 0 at solid sites, which are never read).
// One 32-bit word, written from the device (ordered on the stream and
// graph-safe, unlike a host memset).
static __global__ void k_set_u32(unsigned* p, unsigned v) {
    if (threadIdx.x == 0) *G(p) = v;
}

template <class ST>
__global__ void k_gather(Grid2 g, const typename ST::S* __restrict__ src, typename ST::S* __restrict__ dst) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    const unsigned nb = nbits_at(g, x, y);
    typename ST::T p[Q];
    if (nb & 1u) {
        load_site<ST, true>(src, g.plane, g.pitch, x, wrap1(x - 1, g.nx), wrap1(x + 1, g.nx), y,
                            wrap1(y - 1, g.ny), wrap1(y + 1, g.ny), nb, p);
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) p[i] = typename ST::T(0);
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) ST::st(dst + i * g.plane + s, p[i]);
}

// Storage -> FP64 populations (pitched), with an optional gather (phase P
// -> S).  Deviation forms add w_i*rho0 back; solid sites come out as 0 when
// gathering (the reference zero-fills them in stream).
template <class ST, bool GATHER>
__global__ void k_to_f64(Grid2 g, const typename ST::S* __restrict__ src, double rho0, double* __restrict__ out) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
    const unsigned nb = nbits_at(g, x, y);
    typename ST::T p[Q];
    const bool fl = nb & 1u;
    if (!GATHER || fl) {
        load_site<ST, GATHER>(src, g.plane, g.pitch, x, wrap1(x - 1, g.nx), wrap1(x + 1, g.nx), y,
                              wrap1(y - 1, g.ny), wrap1(y + 1, g.ny), nb, p);
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        double v;
        if (GATHER && !fl) v = 0.0;
        else if (ST::kDev) v = (double)p[i] + weq(i) * rho0;
        else v = (double)p[i];
        out[i * g.plane + s] = v;
    }
}

// FP64 populations -> storage form.
template <class ST>
__global__ void k_from_f64(Grid2 g, const double* __restrict__ in, double rho0, typename ST::S* __restrict__ dst) {
    int x, y, s;
    if (!site_of(g, x, y, s)) return;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const double v = in[i * g.plane + s];
        if (ST::kDev) ST::st(dst + i * g.plane + s, (typename ST::T)(v - weq(i) * rho0));
        else ST::st(dst + i * g.plane + s, (typename ST::T)v);
    }
}

}  // namespace scl
