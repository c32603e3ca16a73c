// step2x.h -- two fused time steps per launch for the two-component
// deviation forms (FP32 arithmetic; FP32 or FP16 storage).
//
// A launch reads P_{n-1} and writes P_{n+1}; the intermediate post-collision
// populations P_n never leave the workgroup, so a population crosses HBM
// once per two steps.  The per-site arithmetic is that of the single-step
// kernel (collide2_dev, the same gather and bounce-back, P_n rounded to the
// storage form before it is gathered again), so the result is bit for bit
// that of two single steps.
//
// Work decomposition.  A workgroup of 256 threads owns one column per thread
// over a strip of 256 columns and walks a band of rows from south to north.
// Consecutive strips overlap by 6 columns: step n is computed on columns
// 1..254 of the strip (it needs the densities of its neighbours), step n+1 on
// columns 2..253, and only columns 3..252 are written, so each strip outputs
// 250 columns and every output column lies inside exactly one strip.  Bands
// recompute two rows of step n on each side.  Per row r of step n:
//   (a) load S_n(r + 1) = Gather(P_{n-1}) from HBM, file its densities (d1);
//   (b) collide row r of step n from d1 rows r-1..r+1 into the ring pr;
//   (c) gather S_{n+1}(r - 1) from pr rows r-2..r, file its densities (d2);
//   (d) collide row r - 2 of step n+1 from d2 rows r-3..r-1, store to HBM.
// Three barriers per row separate the phases; rings of 4 (d1, d2) and 3 (pr)
// rows are never overwritten while a later phase still reads them.
//
// Zero-density rule across the fused steps: step n reads the equilibrium
// velocities of step n-1 from one buffer and writes another, step n+1 reads
// that one and writes a third (CompArgs::ueqx_in / ueqx of a1 and a2).  A
// column computed by two workgroups (the overlap) writes identical values,
// and no buffer is read and written in the same launch except by the
// owning thread in program order.
#pragma once
#include "step.h"

namespace scl {

constexpr int kX2Threads = 256;
constexpr int kX2Halo = 3;
constexpr int kX2Out = kX2Threads - 2 * kX2Halo;   // output columns per strip

template <class ST>
struct X2Lds {
    float d1[4][2][kX2Threads];               // step n: deviation densities of S_n
    float d2[4][2][kX2Threads];               // step n+1: deviation densities of S_{n+1}
    typename ST::S pr[3][2][Q][kX2Threads];   // step n: post-collision populations, as stored
};

// a1: step n (src = P_{n-1}, u_eq buffers of step n); a2: step n+1 (dst =
// P_{n+1}, u_eq buffers of step n+1).  Geometry, rows per band and the
// number of strips (gx) are taken from a1.  Requires nx >= kX2Threads.
template <class ST, bool OPEN>
__global__ __launch_bounds__(kX2Threads) void k_step2x(const StepArgs<float> a1, const StepArgs<float> a2) {
    using S = typename ST::S;
    using OffT = uint32_t;
    constexpr OffT ES = sizeof(S);
    __shared__ X2Lds<ST> L;

    const int nx = a1.nx, ny = a1.ny, wpr = a1.wpr;
    const OffT plane_b = (OffT)a1.plane * ES, pitch_b = (OffT)a1.pitch * ES;
    const uint64_t* __restrict__ mask = a1.mask;
    auto tri = [&](int row, int cm, int c, int cp) -> unsigned {
        if constexpr (OPEN) return 7u;
        else return triple(mask, wpr, row, cm, c, cp);
    };
    auto neigh = [&](unsigned s, unsigned c, unsigned n) -> unsigned {
        if constexpr (OPEN) return 0x1FFu;
        else return neigh_bits(s, c, n);
    };

    // Workgroup -> (strip, band), XCD-aware as in k_step.
    const int nbx = a1.gx;
    const int nby = (ny + a1.rows - 1) / a1.rows;
    int bid = blockIdx.x;
    {
        const int nblk = nbx * nby;
        const int xcd = bid & 7, k = bid >> 3;
        const int per = (nblk + 7) >> 3;
        const int full = nblk - 8 * (per - 1);
        bid = xcd < full ? xcd * per + k : full * per + (xcd - full) * (per - 1) + k;
    }
    const int bx = bid % nbx, by = bid / nbx;

    const int t = threadIdx.x;
    const int x0 = bx * kX2Out;
    int X = x0 - kX2Halo + t;   // in [-3, nx + 251]: one wrap at most (nx >= 256)
    X = X < 0 ? X + nx : (X >= nx ? X - nx : X);
    const int Xm = wrap1(X - 1, nx), Xp = wrap1(X + 1, nx);
    const int nout = min(kX2Out, nx - x0);
    const bool out_lane = t >= kX2Halo && t < kX2Halo + nout;
    const bool s1_lane = t >= 1 && t < kX2Threads - 1;   // step n valid
    const bool s2_lane = t >= 2 && t < kX2Threads - 2;   // step n+1 gather valid
    const int y0 = by * a1.rows, y1 = min(y0 + a1.rows, ny);
    auto wr = [&](int r) {
        r %= ny;
        return r < 0 ? r + ny : r;
    };
    auto r4 = [&](int r) { return (r - y0 + 8) & 3; };   // r >= y0 - 3
    auto r3 = [&](int r) { return (r - y0 + 9) % 3; };

    auto nbits = [&](int r) -> unsigned {
        return neigh(tri(wr(r - 1), Xm, X, Xp), tri(wr(r), Xm, X, Xp), tri(wr(r + 1), Xm, X, Xp));
    };
    // S_n(X, r) = Gather(P_{n-1}), zero at solid sites
    auto gather1 = [&](int r, unsigned nb, float (&p)[2][Q]) {
        OffT off[Q];
        pull_offsets<ST, true, OffT>(plane_b, pitch_b, X, Xm, Xp, wr(r), wr(r - 1), wr(r + 1), nb, off);
        if (nb & 1u) {
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int i = 0; i < Q; ++i) p[c][i] = ld_at<ST, OffT>(a1.c[c].src, off[i]);
        } else {
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int i = 0; i < Q; ++i) p[c][i] = 0.f;
        }
    };
    // neighbour densities of (t, r) from a density ring
    auto nbr_dens = [&](float (&ring)[4][2][kX2Threads], int r, float (&nr)[2][Q], float (&rc)[2]) {
        const int sS = r4(r - 1), sC = r4(r), sN = r4(r + 1);
#pragma unroll
        for (int i = 1; i < Q; ++i) {
            const int sl = cy(i) == 1 ? sN : (cy(i) == -1 ? sS : sC);
#pragma unroll
            for (int c = 0; c < 2; ++c) nr[c][i] = ring[sl][c][t + cx(i)];
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            rc[c] = ring[sC][c][t];
            nr[c][0] = rc[c];
        }
    };

    float pA[2][Q], pB[2][Q];     // S_n rows r (collided now) and r + 1 (loaded now)
    float s2c[2][Q], s2p[2][Q];   // S_{n+1} rows r - 1 (gathered now) and r - 2 (collided now)
    unsigned nb_p1 = 0, nb_0, nb_m1 = 0, nb_m2 = 0;   // neighbourhoods of rows r+1, r, r-1, r-2

    // prologue: densities of S_n rows y0-3, y0-2; populations of row y0-2
    {
        const unsigned nbq = nbits(y0 - 3);
        gather1(y0 - 3, nbq, pB);
#pragma unroll
        for (int c = 0; c < 2; ++c) L.d1[r4(y0 - 3)][c][t] = dens(pB[c]);
        nb_0 = nbits(y0 - 2);
        gather1(y0 - 2, nb_0, pA);
#pragma unroll
        for (int c = 0; c < 2; ++c) L.d1[r4(y0 - 2)][c][t] = dens(pA[c]);
    }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int i = 0; i < Q; ++i) s2p[c][i] = s2c[c][i] = 0.f;

    for (int r = y0 - 2; r <= y1 + 1; ++r) {
        // (a) step n: S_n(r + 1)
        nb_p1 = nbits(r + 1);
        gather1(r + 1, nb_p1, pB);
#pragma unroll
        for (int c = 0; c < 2; ++c) L.d1[r4(r + 1)][c][t] = dens(pB[c]);
        __syncthreads();
        // (b) step n: collide row r into the ring
        if (s1_lane && (OPEN || (nb_0 & 1u))) {
            float nr[2][Q], rc[2], oa[Q], ob[Q];
            nbr_dens(L.d1, r, nr, rc);
            collide2_dev(a1, wr(r) * a1.pitch + X, pA, rc, nr, OPEN ? 0x1FFu : nb_0, oa, ob);
            const int sl = r3(r);
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                L.pr[sl][0][i][t] = ST::enc(oa[i]);
                L.pr[sl][1][i][t] = ST::enc(ob[i]);
            }
        }
        __syncthreads();
        // (c) step n+1: gather S_{n+1}(r - 1) from the ring
        if (r >= y0 && s2_lane) {
            const int g = r - 1;
            if (OPEN || (nb_m1 & 1u)) {
                const int sC = r3(g), sS = r3(g - 1), sN = r3(g + 1);
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    s2c[c][0] = ST::dec(L.pr[sC][c][0][t]);
#pragma unroll
                    for (int i = 1; i < Q; ++i) {
                        // source x - c_i: row g - cy(i), column t - cx(i)
                        const int ss = cy(i) == 1 ? sS : (cy(i) == -1 ? sN : sC);
                        const bool src_fluid = OPEN || ((nb_m1 >> opp(i)) & 1u);
                        const S v = src_fluid ? L.pr[ss][c][i][t - cx(i)] : L.pr[sC][c][opp(i)][t];
                        s2c[c][i] = ST::dec(v);
                    }
                }
            } else {
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int i = 0; i < Q; ++i) s2c[c][i] = 0.f;
            }
#pragma unroll
            for (int c = 0; c < 2; ++c) L.d2[r4(g)][c][t] = dens(s2c[c]);
        }
        __syncthreads();
        // (d) step n+1: collide row r - 2, store P_{n+1}
        const int h = r - 2;
        if (h >= y0 && h < y1 && out_lane && (OPEN || (nb_m2 & 1u))) {
            float nr[2][Q], rc[2], oa[Q], ob[Q];
            nbr_dens(L.d2, h, nr, rc);
            collide2_dev(a2, h * a2.pitch + X, s2p, rc, nr, OPEN ? 0x1FFu : nb_m2, oa, ob);
            const OffT home = (OffT)h * pitch_b + (OffT)X * ES;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                st_at<ST, OffT>(a2.c[0].dst, (OffT)i * plane_b + home, oa[i]);
                st_at<ST, OffT>(a2.c[1].dst, (OffT)i * plane_b + home, ob[i]);
            }
        }
        // rotate
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                s2p[c][i] = s2c[c][i];
                pA[c][i] = pB[c][i];
            }
        nb_m2 = nb_m1;
        nb_m1 = nb_0;
        nb_0 = nb_p1;
    }
}

}  // namespace scl
