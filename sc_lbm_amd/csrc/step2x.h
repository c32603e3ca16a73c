// step2x.h -- two fused time steps per launch for the two-component
// deviation forms (FP32 arithmetic; FP32 or FP16 storage).
//
// step2x_tile is the two-step tile of the server's persistent kernel, whose
// per-job arguments carry separate equilibrium-velocity pointers.  Fields
// launch k_step2x (stepk.h), the K = 2 case of the K-step tile, with the
// same strips, the same LDS layout and the same results.
//
// A launch reads P_{n-1} and writes P_{n+1}; the intermediate post-collision
// populations P_n never leave the workgroup, so a population crosses HBM
// once per two steps.  The per-site arithmetic is that of the single-step
// kernel (collide2_core, the same gather and bounce-back, P_n rounded to the
// storage form before it is gathered again), so the result is bit for bit
// that of two single steps.  Both collisions take their physics parameters
// from a1 (a2 holds the same ones; reading one set keeps the scalar
// registers from spilling) and their equilibrium-velocity buffers apart.
//
// Work decomposition.  A workgroup of 256 threads owns one column per thread
// over a strip of 256 columns and walks a band of rows from south to north.
// Step n is valid on columns 1..254 of the strip (it needs the densities of
// its neighbours), step n+1 on columns 2..253; a strip writes columns
// 3..3+OUT-1 (OUT = x2_out, at most 250) and the next strip starts OUT
// columns further, so every output column lies inside exactly one strip.
// Bands recompute two rows of step n on each side.  Per row r of step n:
//   (a) file the densities of S_n(r + 1) = Gather(P_{n-1}) (d1), whose loads
//       were issued one row earlier, and issue the loads of S_n(r + 2) (they
//       stay in flight through (b)-(d): the barriers wait only for LDS);
//   (b) collide row r of step n from d1 rows r-1..r+1 into the ring pr;
//   (c) gather S_{n+1}(r - 1) from pr rows r-2..r, file its densities (d2);
//   (d) collide row r - 2 of step n+1 from d2 rows r-3..r-1, store to HBM.
// Three barriers per row separate the phases.  LDS holds only what a later
// phase still reads: densities in rings of 3 rows; post-collision
// populations by direction class (phase (c) at row r reads direction i from
// row r - 1 - c_iy): c_iy = +1 in 3 rows, c_iy = 0 in 2, c_iy = -1 in 1 (2
// with solids, whose bounce-back reads the own row r - 1 in every
// direction); the rest population stays in a register.  That fits three
// workgroups per CU with FP32 storage.
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

// Threads (columns) per strip.  Wider strips re-read fewer halo columns but
// synchronise more waves per barrier: at 4096^2 FP32, 384 / 448 / 512
// threads ran at 43.7k / 49.4k / 49.6k MLUPS against 57.6k with 256.
#ifndef SCL_X2_THREADS
#define SCL_X2_THREADS 256
#endif
constexpr int kX2Threads = SCL_X2_THREADS;
constexpr int kX2Halo = 3;
// Output columns per strip: with FP32 storage the threads minus 32, i.e.
// 224 (seven whole 128-byte lines per row and plane, so no line is written
// partly by two strips); with FP16 storage every valid column, 250 (the
// aligned 192 would cost more arithmetic than the split lines cost in
// traffic).  Stores are cached rather than streamed: the two halves of a
// split line then merge in L2.
template <class ST>
__host__ __device__ constexpr int x2_out() {
    return sizeof(typename ST::S) == 4 ? kX2Threads - 32 : kX2Threads - 2 * kX2Halo;
}

// direction class (0: c_y = +1, 1: c_y = 0, 2: c_y = -1) and index within it
__host__ __device__ constexpr int x2_class(int i) { return cy(i) == 1 ? 0 : (cy(i) == 0 ? 1 : 2); }
__host__ __device__ constexpr int x2_index(int i) { return i == 1 || i == 2 || i == 4 ? 0 : (i == 3 || i == 5 || i == 7 ? 1 : 2); }

template <class ST, bool OPEN>
struct X2Lds {
    float d1[3][2][kX2Threads];                      // step n: deviation densities of S_n
    float d2[3][2][kX2Threads];                      // step n+1: deviation densities of S_{n+1}
    // step n: post-collision populations as stored, by direction class
    typename ST::S pu[3][2][3][kX2Threads];          // c_y = +1 (2, 5, 6)
    typename ST::S pz[2][2][2][kX2Threads];          // c_y = 0 (1, 3)
    typename ST::S pd[OPEN ? 1 : 2][2][3][kX2Threads];   // c_y = -1 (4, 7, 8)
};

// One tile: strip bx (columns bx*OUT-3 .. bx*OUT+252), band of rows
// [y0, y1).  a1: step n (src = P_{n-1}, u_eq buffers of step n); a2: step
// n+1 (dst = P_{n+1}, u_eq buffers of step n+1); geometry from a1.
// Requires nx >= kX2Threads; the caller separates consecutive tiles of a
// workgroup with a barrier (L is reused).
template <class ST, bool OPEN, bool FORCE = true>
__device__ inline void step2x_tile(const StepArgs<float>& a1, const StepArgs<float>& a2, const int bx, const int y0,
                                   const int y1, X2Lds<ST, OPEN>& L) {
    using S = typename ST::S;
    using OffT = uint32_t;
    constexpr OffT ES = sizeof(S);

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

    const UeqPtrs u1 = ueq_ptrs(a1), u2 = ueq_ptrs(a2);
    const int t = threadIdx.x;
    constexpr int OUT = x2_out<ST>();
    const int x0 = bx * OUT;
    int X = x0 - kX2Halo + t;   // in [-3, nx + 251]: one wrap at most (nx >= 256)
    X = X < 0 ? X + nx : (X >= nx ? X - nx : X);
    const int Xm = wrap1(X - 1, nx), Xp = wrap1(X + 1, nx);
    const int nout = min(OUT, nx - x0);
    const bool out_lane = t >= kX2Halo && t < kX2Halo + nout;
    const bool s1_lane = t >= 1 && t < kX2Threads - 1;   // step n valid
    const bool s2_lane = t >= 2 && t < kX2Threads - 2;   // step n+1 gather valid
    auto wr = [&](int r) {
        r %= ny;
        return r < 0 ? r + ny : r;
    };
    auto r3 = [&](int r) { return (r - y0 + 9) % 3; };   // r >= y0 - 9
    auto r2 = [&](int r) { return (r - y0 + 4) & 1; };   // r >= y0 - 4
    // LDS slot of direction i (1..8) of row `row`, column `col`
    auto pr = [&](int row, int c, int i, int col) -> S& {
        if (x2_class(i) == 0) return L.pu[r3(row)][c][x2_index(i)][col];
        if (x2_class(i) == 1) return L.pz[r2(row)][c][x2_index(i)][col];
        return L.pd[OPEN ? 0 : r2(row)][c][x2_index(i)][col];
    };

    auto nbits = [&](int r) -> unsigned {
        return neigh(tri(wr(r - 1), Xm, X, Xp), tri(wr(r), Xm, X, Xp), tri(wr(r + 1), Xm, X, Xp));
    };
    // S_n(X, r) = Gather(P_{n-1}), zero at solid sites
    // A solid site loads its own cell (plane 0) in every direction and keeps
    // zeros: no branch round the loads, which would make the compiler wait
    // at the join for every access in flight.
    auto gather1 = [&](int r, unsigned nb, float (&p)[2][Q]) {
        OffT off[Q];
        pull_offsets<ST, true, OffT>(plane_b, pitch_b, X, Xm, Xp, wr(r), wr(r - 1), wr(r + 1), nb, off);
        const bool fluid = OPEN || (nb & 1u);
        const OffT home = (OffT)wr(r) * pitch_b + (OffT)X * ES;
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const float v = ld_at<ST, OffT>(a1.c[c].src, fluid ? off[i] : home);
                p[c][i] = fluid ? v : 0.f;
            }
    };
    // Mask rows, pipelined like the populations (with solids): the words of
    // row r + 4 load during row r; each row's triple is formed once.
    uint64_t mw[3] = {0, 0, 0};
    auto mask_words = [&](int row) {
        if constexpr (!OPEN) {
            const size_t base = (size_t)wr(row) * wpr;
            mw[0] = G(mask)[base + (Xm >> 6)];
            mw[1] = G(mask)[base + (X >> 6)];
            mw[2] = G(mask)[base + (Xp >> 6)];
        }
    };
    auto tri_of_words = [&]() -> unsigned {
        if constexpr (OPEN) return 7u;
        else
            return (unsigned)((mw[0] >> (Xm & 63)) & 1ull) | ((unsigned)((mw[1] >> (X & 63)) & 1ull) << 1) |
                   ((unsigned)((mw[2] >> (Xp & 63)) & 1ull) << 2);
    };
    // neighbour densities of (t, r) from a density ring
    auto nbr_dens = [&](float (&ring)[3][2][kX2Threads], int r, float (&nr)[2][Q], float (&rc)[2]) {
        const int sS = r3(r - 1), sC = r3(r), sN = r3(r + 1);
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

    // S_n populations rotate by name through three arrays (the loop below is
    // unrolled by three): a copy of an array whose loads are in flight would
    // make the compiler wait for every outstanding access, stores included.
    float pX[2][Q], pY[2][Q], pZ[2][Q];
    float s2c[2][Q], s2p[2][Q];   // S_{n+1} rows r - 1 (gathered now) and r - 2 (collided now)
    unsigned nb_p2 = 0, nb_p1, nb_0, nb_m1 = 0, nb_m2 = 0;   // neighbourhoods of rows r+2 .. r-2

    // prologue: densities of S_n rows y0-3, y0-2; populations of row y0-2 (pX);
    // loads of row y0-1 (pY) in flight
    {
        const unsigned nbq = nbits(y0 - 3);
        gather1(y0 - 3, nbq, pZ);
#pragma unroll
        for (int c = 0; c < 2; ++c) L.d1[r3(y0 - 3)][c][t] = dens(pZ[c]);
        nb_0 = nbits(y0 - 2);
        gather1(y0 - 2, nb_0, pX);
#pragma unroll
        for (int c = 0; c < 2; ++c) L.d1[r3(y0 - 2)][c][t] = dens(pX[c]);
        nb_p1 = nbits(y0 - 1);
        gather1(y0 - 1, nb_p1, pY);
        mask_words(y0 + 1);
    }
    unsigned t_r1 = tri(wr(y0 - 1), Xm, X, Xp), t_r2 = tri(wr(y0), Xm, X, Xp);   // triples of rows r+1, r+2
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int i = 0; i < Q; ++i) s2p[c][i] = s2c[c][i] = 0.f;
    float rest_now[2] = {0.f, 0.f}, rest_prev[2] = {0.f, 0.f};   // rest populations of rows r, r - 1

    // one row r: pa = S_n row r (densities filed), pb = row r + 1 (loads issued
    // one row earlier), pn receives the loads of row r + 2
    auto row = [&](const int r, float (&pa)[2][Q], float (&pb)[2][Q], float (&pn)[2][Q]) {
        // (a) step n: loads of S_n(r + 2); densities of S_n(r + 1).  The last
        // row (r = y1 + 1) reloads row y1 + 2 (just read, so cached) instead
        // of branching round the loads: after a branch the compiler could no
        // longer wait for pb's loads alone.
        const unsigned t_r3 = tri_of_words();   // row r + 3
        mask_words(r + 4);
        nb_p2 = neigh(t_r1, t_r2, t_r3);
        t_r1 = t_r2;
        t_r2 = t_r3;
        gather1(min(r + 2, y1 + 2), nb_p2, pn);
#pragma unroll
        for (int c = 0; c < 2; ++c) L.d1[r3(r + 1)][c][t] = dens(pb[c]);
        __syncthreads();
        // (b) step n: collide row r into the ring
        if (s1_lane && (OPEN || (nb_0 & 1u))) {
            float nr[2][Q], rc[2], oa[Q], ob[Q];
            nbr_dens(L.d1, r, nr, rc);
            collide2_core<float, UeqPtrs, FORCE>(a1, u1, wr(r) * a1.pitch + X, pa, rc, nr, OPEN ? 0x1FFu : nb_0, oa, ob);
#pragma unroll
            for (int i = 1; i < Q; ++i) {
                pr(r, 0, i, t) = ST::enc(oa[i]);
                pr(r, 1, i, t) = ST::enc(ob[i]);
            }
            rest_now[0] = ST::dec(ST::enc(oa[0]));
            rest_now[1] = ST::dec(ST::enc(ob[0]));
        }
        __syncthreads();
        // (c) step n+1: gather S_{n+1}(r - 1) from the ring
        if (r >= y0 && s2_lane) {
            const int g = r - 1;
            if (OPEN || (nb_m1 & 1u)) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    s2c[c][0] = rest_prev[c];
#pragma unroll
                    for (int i = 1; i < Q; ++i) {
                        // source x - c_i: row g - cy(i), column t - cx(i)
                        const bool src_fluid = OPEN || ((nb_m1 >> opp(i)) & 1u);
                        const S v = src_fluid ? pr(g - cy(i), c, i, t - cx(i)) : pr(g, c, opp(i), t);
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
            for (int c = 0; c < 2; ++c) L.d2[r3(g)][c][t] = dens(s2c[c]);
        }
        __syncthreads();
        // (d) step n+1: collide row r - 2, store P_{n+1}
        const int h = r - 2;
        if (h >= y0 && h < y1 && out_lane && (OPEN || (nb_m2 & 1u))) {
            float nr[2][Q], rc[2], oa[Q], ob[Q];
            nbr_dens(L.d2, h, nr, rc);
            // step n+1: the parameters of a1 (the same), the buffers of a2
            collide2_core<float, UeqPtrs, FORCE>(a1, u2, h * a1.pitch + X, s2p, rc, nr, OPEN ? 0x1FFu : nb_m2, oa, ob);
            const OffT home = (OffT)h * pitch_b + (OffT)X * ES;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                ST::stc((S*)((char*)a2.c[0].dst + (size_t)i * plane_b + home), oa[i]);
                ST::stc((S*)((char*)a2.c[1].dst + (size_t)i * plane_b + home), ob[i]);
            }
        }
        // rotate (the S_n populations rotate by name)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
#pragma unroll
            for (int i = 0; i < Q; ++i) s2p[c][i] = s2c[c][i];
            rest_prev[c] = rest_now[c];
        }
        nb_m2 = nb_m1;
        nb_m1 = nb_0;
        nb_0 = nb_p1;
        nb_p1 = nb_p2;
    };
    // rows y0 - 2 .. y1 + 1 (at least three); the bounds are uniform over the
    // workgroup, so every thread reaches the same barriers
    for (int r = y0 - 2;; r += 3) {
        row(r, pX, pY, pZ);
        if (r + 1 > y1 + 1) break;
        row(r + 1, pY, pZ, pX);
        if (r + 2 > y1 + 1) break;
        row(r + 2, pZ, pX, pY);
        if (r + 3 > y1 + 1) break;
    }
}

}  // namespace scl
