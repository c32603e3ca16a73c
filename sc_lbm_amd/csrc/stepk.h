// stepk.h -- K fused time steps per launch for the two-component deviation
// forms (FP32 arithmetic; FP32 or FP16 storage): K = 2 (k_step2x, and the
// server's k_serve2x) and K = 3 (k_stepk).
//
// A launch reads P_{n-1} and writes P_{n+K-1}; the post-collision
// populations of the K-1 intermediate steps stay in LDS, so a population
// crosses HBM once per K steps.  Every stage computes a site exactly as a
// single step does (collide2_core, the same gather and bounce-back, each
// intermediate rounded to the storage form before it is gathered again), so
// the result is bit for bit that of K single steps.
//
// Work decomposition.  A workgroup of T threads owns one column per thread
// and walks a band of rows from south to north.  While stage 1 collides row
// r, stage s (1..K) collides row r - 2(s-1) (per-stage schedule) or
// r - 3(s-1) (two-phase schedule), having gathered it from the
// post-collision ring of stage s-1 in the iteration before.  Stage s is valid on
// columns 2s-1 .. T-2s of the strip, so a strip writes OUT columns from
// column H = 2K-1 on (OUT: a multiple of 32 with FP32 storage, so no
// 128-byte line is written by two strips) and bands recompute 2(K-1) rows
// of the first stage on each side.  Per row: file the densities of
// S_n(r + 1), loaded one row earlier, and issue the loads of row r + 2; then
// each stage's gather and collision, in one of two schedules (two_phase
// below).  The barriers wait only for LDS, so the loads stay in flight.
//
// LDS per stage: densities in rings of 3 rows; post-collision populations
// by direction class (the next stage's gather at row g reads direction i
// from row g - c_iy): c_iy = +1 in 3 rows, c_iy = 0 in 2, c_iy = -1 in 1 (2
// with solids, whose bounce-back reads the own row in every direction); the
// rest population stays in a register.
//
// Zero-density rule: stage s reads the equilibrium velocities of the step
// before from one buffer set and writes the next, which stage s + 1 reads;
// the K + 1 sets of a launch are distinct (K < the ring's set count), so no
// buffer is read by one workgroup while another writes it (columns computed
// by two strips write identical values).
#pragma once
#include "step2x.h"

namespace scl {

// Fewest lattice rows a K-step tile takes: its rows (halos and mask rows
// included) reach 2K + 1 rows past the lattice on either side, and a single
// conditional add or subtract must wrap them.
template <int K>
__host__ __device__ constexpr int xk_min_rows() {
    return 2 * K + 2;
}

template <class ST, int K, int T>
__host__ __device__ constexpr int xk_out() {
    return sizeof(typename ST::S) == 4 ? (T - 2 * (2 * K - 1)) / 32 * 32 : T - 2 * (2 * K - 1);
}

// Threads per strip of the three-step kernel: the widest strip (at most
// 256) whose LDS lets two workgroups share a CU (160 KiB).  LDS per column:
// FP32 storage 328 B solid-free, 376 B with solids; FP16 storage 200 / 224 B.
template <class ST, bool OPEN>
__host__ __device__ constexpr int x3_threads() {
    return sizeof(typename ST::S) == 2 ? 256 : (OPEN ? 240 : 208);
}

template <class ST, bool OPEN, int K, int T>
struct XkLds {
    typename ST::T d[K][3][2][T];                        // densities of stage s
    typename ST::S pu[K - 1][3][2][3][T];                // post-collision, c_y = +1 (2, 5, 6)
    typename ST::S pz[K - 1][2][2][2][T];                // c_y = 0 (1, 3)
    typename ST::S pd[K - 1][OPEN ? 1 : 2][2][3][T];     // c_y = -1 (4, 7, 8)
};

// Equilibrium-velocity buffers of a deviation-form field: nsets sets of
// (x, y) buffers of n doubles each, component c's set k direction d at
// base[c] + k * set_stride + d * n.  A step reads set cur and writes set
// cur + 1 (mod nsets).  A few scalar registers describe every set of both
// components, where separate pointers for K stages would take 16 K and
// spill.  Fields hold kUeqRing sets (one more than the most steps a launch
// fuses), the server three.
constexpr int kUeqRing = 4;
struct UeqRing {
    double* base[2];
    long long n, set_stride;
    int nsets;
    int cur[2];   // < nsets
};
// The buffers stage s (0-based) of a launch reads and writes.
struct UeqStage {
    const UeqRing& r;
    int s;
    __device__ double* buf(int c, int next, int d) const {
        int k = r.cur[c] + s + next;   // < 2 nsets: cur < nsets, s + next <= K < nsets
        k -= k >= r.nsets ? r.nsets : 0;
        return r.base[c] + k * r.set_stride + d * r.n;
    }
    __device__ const double* ix(int c) const { return buf(c, 0, 0); }
    __device__ const double* iy(int c) const { return buf(c, 0, 1); }
    __device__ double* ox(int c) const { return buf(c, 1, 0); }
    __device__ double* oy(int c) const { return buf(c, 1, 1); }
};

// Arguments of one K-step launch: physics, geometry and P_{n-1} (a.c[c].src)
// from `a` (its equilibrium-velocity pointers are unused); P_{n+K-1} into
// dst; the equilibrium-velocity sets in u (deviation forms).  FP64: the
// launch's number (nonzero) and the word a site below the zero-density
// rule's watch threshold sets to it (collide2_f64); a launch that finds the
// word set returns at once.
template <int K, class R = float>
struct XkArgs {
    StepArgs<R> a;
    void* dst[2];
    UeqRing u;
    unsigned* flag;
    unsigned launch;
    double watch;   // >= kUeqWatch64
};

// The row schedule of a tile.  Two phases (stage s lags stage 1 by 3(s - 1)
// rows, two barriers per row) where measured faster: FP32 storage, no
// solids, three steps (76.7k against 72.8k MLUPS at 4096^2), and FP64
// without solids (two steps: 1312 -> 1257 us per launch at 4096^2; one
// wave per SIMD, held there by LDS, so its registers are spare).  Elsewhere
// the stages follow each other within the row (2K - 1 barriers): there the
// two-phase schedule's longer register lifetimes cost occupancy (FP16 two
// steps: 72.4k against 82.9k; FP32 three steps with solids: 26.0k against
// 41.7k).  Both give the same bits.
template <class ST, bool OPEN, int K>
__host__ __device__ constexpr bool two_phase() {
    return OPEN && ((sizeof(typename ST::S) == 4 && K == 3) || sizeof(typename ST::S) == 8);
}

// Whether the row loop is unrolled by 6, so that the two register arrays of
// stages s >= 2 (the row gathered for the next iteration, the row collided
// now) swap roles by name, or by 3 with a copy per row.  Renaming saves 6-8%
// of the vector instructions; it is kept where it measured faster or equal,
// the three-step tiles (solid-free FP32 4096^2: 76.8k -> 78.0k MLUPS, FP16:
// 74.3k -> 75.1k; porous FP32 equal, FP16 45.1k -> 46.6k).  With two steps
// it gained nothing (FP32) or cost occupancy (FP16: 164 -> 177 VGPRs, three
// waves per SIMD to two).
template <class ST, bool OPEN, int K>
__host__ __device__ constexpr bool rename_rows() {
    return K == 3;
}

// Whether the three-row ring slots are compile-time constants (the LDS
// addresses then need no per-row scalar slot arithmetic: 20-37% fewer
// scalar and 2.5-3.7% fewer vector instructions).  With solids only
// together with fresh_lane (alone, the FP32 three-step tile went past 256
// VGPRs).
template <class ST, bool OPEN, int K>
__host__ __device__ constexpr bool const_slots() {
    return OPEN ? (K == 3 || sizeof(typename ST::S) == 2) : true;
}

// Whether each row takes the lane index afresh (an empty asm that the
// compiler must assume changes it), so that the LDS addresses derived from
// it are formed inside the row rather than hoisted out of the loop, each in
// a register of its own.  FP16 two steps: 164 -> 124 VGPRs, four waves per
// SIMD instead of three; with solids it keeps the compile-time slots within
// the registers of the same occupancy.  The solid-free FP32 tiles are held
// to their occupancy by LDS, where the hoisted addresses save instructions.
template <class ST, bool OPEN, int K>
__host__ __device__ constexpr bool fresh_lane() {
    return OPEN ? (K == 2 && sizeof(typename ST::S) == 2) : true;
}

// A row index and its ring phase: (r - the tile's first row) modulo 3 (6
// with the loop unrolled by 6), a constant in every unrolled call.
struct RowAnchor {
    int r, ph;
};
template <int P>
using Ph = std::integral_constant<int, P>;

// Calls row(phase, r, ...) for r = r_first .. r_last, the populations of S_n
// rotating by name through pX, pY, pZ.  The bounds are uniform over the
// workgroup, so every thread reaches the same barriers.  The row index is
// read from lane 0 to tell the compiler so: otherwise it keeps the loop
// counter and every row computation in vector registers (a divergent loop).
template <bool RENAME, class Row, class P, class G>
__device__ __attribute__((always_inline)) inline void walk_rows(const int r_first, const int r_last, Row& row, P& pX,
                                                                P& pY, P& pZ, G& sA, G& sB) {
    if constexpr (RENAME) {
        for (int r0 = r_first;; r0 += 6) {
            const int r = __builtin_amdgcn_readfirstlane(r0);
            row(Ph<0>{}, r, pX, pY, pZ, sA, sB);
            if (r + 1 > r_last) break;
            row(Ph<1>{}, r + 1, pY, pZ, pX, sB, sA);
            if (r + 2 > r_last) break;
            row(Ph<2>{}, r + 2, pZ, pX, pY, sA, sB);
            if (r + 3 > r_last) break;
            row(Ph<3>{}, r + 3, pX, pY, pZ, sB, sA);
            if (r + 4 > r_last) break;
            row(Ph<4>{}, r + 4, pY, pZ, pX, sA, sB);
            if (r + 5 > r_last) break;
            row(Ph<5>{}, r + 5, pZ, pX, pY, sB, sA);
            if (r + 6 > r_last) break;
        }
    } else {   // row() copies sA to sB at its end
        for (int r0 = r_first;; r0 += 3) {
            const int r = __builtin_amdgcn_readfirstlane(r0);
            row(Ph<0>{}, r, pX, pY, pZ, sA, sB);
            if (r + 1 > r_last) break;
            row(Ph<1>{}, r + 1, pY, pZ, pX, sA, sB);
            if (r + 2 > r_last) break;
            row(Ph<2>{}, r + 2, pZ, pX, pY, sA, sB);
            if (r + 3 > r_last) break;
        }
    }
}

template <class ST, bool OPEN, int K, int T, bool FORCE>
__device__ inline void stepk_tile(const XkArgs<K, typename ST::T>& A, const int bx, const int y0, const int y1,
                                  XkLds<ST, OPEN, K, T>& L) {
    using S = typename ST::S;
    using R = typename ST::T;   // arithmetic: float (deviation forms) or double
    using OffT = uint32_t;
    constexpr OffT ES = sizeof(S);
    constexpr int H = 2 * K - 1;            // halo columns on each side
    constexpr int OUT = xk_out<ST, K, T>();
    const StepArgs<R>& a = A.a;
    const int nx = a.nx, ny = a.ny, wpr = a.wpr;
    const OffT plane_b = (OffT)a.plane * ES, pitch_b = (OffT)a.pitch * ES;
    const uint64_t* __restrict__ mask = a.mask;
    auto tri = [&](int row, int cm, int c, int cp) -> unsigned {
        if constexpr (OPEN) return 7u;
        else return triple(mask, wpr, row, cm, c, cp);
    };
    auto neigh = [&](unsigned s, unsigned c, unsigned n) -> unsigned {
        if constexpr (OPEN) return 0x1FFu;
        else return neigh_bits(s, c, n);
    };

    int t = threadIdx.x;   // (fresh_lane)
    const int x0 = bx * OUT;
    const int nout = min(OUT, nx - x0);
    const bool out_lane = t >= H && t < H + nout;
    // Lanes the strip needs: its outputs and H halo columns on each side.
    // The lanes past them (a few in every FP32 strip, whole waves in a
    // lattice's narrow last strip) collide nothing: every stage's lane bounds
    // below end at tl instead of T.  They load the last needed lane's column,
    // so their loads fall in that lane's cache lines and cost no traffic.
    const int tl = min(T, nout + 2 * H);
    int X = x0 - H + min(t, tl - 1);   // one wrap at most (nx >= T)
    X = X < 0 ? X + nx : (X >= nx ? X - nx : X);
    const int Xm = wrap1(X - 1, nx), Xp = wrap1(X + 1, nx);
    // Periodic row index.  Every row a tile touches (mask rows included) lies
    // within 2K + 1 rows of the lattice, and the lattice has at least
    // xk_min_rows<K>() = 2K + 2 rows (the host runs single steps on shorter
    // ones), so one conditional add or subtract wraps it.  (A remainder
    // costs 17 scalar instructions, three row indices a row.)
    auto wr = [&](int r) {
        r = r < 0 ? r + ny : r;
        return r >= ny ? r - ny : r;
    };
    // Ring slots.  Row `row` sits in slot (row - ra) mod 3 of a three-row
    // ring and (row - ra) mod 2 of a two-row ring, ra the first row stage 1
    // collides (row >= ra - 1).  Every row index is an anchor row plus a
    // constant, and the anchor's phase, (an.r - ra) mod 3, is a compile-time
    // constant: with const_slots the three-row slots fold into the LDS
    // offsets.  The two-row slots stay computed (folding them too raised the
    // register count past two waves per SIMD).
    const int ra = y0 - 2 * K + 2;
    auto r3 = [&](const RowAnchor& an, int row) {
        if constexpr (const_slots<ST, OPEN, K>()) return ((an.ph + (row - an.r)) % 3 + 3) % 3;
        else return (row - ra + 3) % 3;
    };
    auto r2 = [&](const RowAnchor&, int row) { return (row - ra) & 1; };
    // LDS slot of direction i (1..8) of row `row`, column `col`, ring of stage q (0-based)
    auto pr = [&](const RowAnchor& an, int q, int row, int c, int i, int col) -> S& {
        if (x2_class(i) == 0) return L.pu[q][r3(an, row)][c][x2_index(i)][col];
        if (x2_class(i) == 1) return L.pz[q][r2(an, row)][c][x2_index(i)][col];
        return L.pd[q][OPEN ? 0 : r2(an, row)][c][x2_index(i)][col];
    };
    auto nbits = [&](int r) -> unsigned {
        return neigh(tri(wr(r - 1), Xm, X, Xp), tri(wr(r), Xm, X, Xp), tri(wr(r + 1), Xm, X, Xp));
    };
    // S_n(X, r) = Gather(P_{n-1}), zero at solid sites (a solid site loads its
    // own cell and keeps zeros: no branch round the loads)
    auto gather1 = [&](int r, unsigned nb, R (&p)[2][Q]) {
        OffT off[Q];
        pull_offsets<ST, true, OffT>(plane_b, pitch_b, X, Xm, Xp, wr(r), wr(r - 1), wr(r + 1), nb, off);
        const bool fluid = OPEN || (nb & 1u);
        const OffT home = (OffT)wr(r) * pitch_b + (OffT)X * ES;
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const R v = ldw_at<ST, OffT>(a.c[c].src, fluid ? off[i] : home);   // working units
                p[c][i] = fluid ? v : R(0);
            }
    };
    uint64_t mw[3] = {0, 0, 0};   // mask words of row r + 4, loading during row r
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
    auto nbr_dens = [&](const RowAnchor& an, int q, int r, R (&nr)[2][Q], R (&rc)[2]) {
        const int sS = r3(an, r - 1), sC = r3(an, r), sN = r3(an, r + 1);
#pragma unroll
        for (int i = 1; i < Q; ++i) {
            const int sl = cy(i) == 1 ? sN : (cy(i) == -1 ? sS : sC);
#pragma unroll
            for (int c = 0; c < 2; ++c) nr[c][i] = L.d[q][sl][c][t + cx(i)];
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            rc[c] = L.d[q][sC][c][t];
            nr[c][0] = rc[c];
        }
    };

    // One fluid site's collision in stage q (0-based): the deviation forms'
    // collide2_core, or the reference's FP64 arithmetic (collide2_f64).
    // oa / ob: the outputs; ea / eb: the same as the storage form encodes them
    auto collide = [&](int q, int s_idx, const R (&p)[2][Q], const R (&rc)[2], const R (&nr)[2][Q], unsigned nb,
                       R (&oa)[Q], R (&ob)[Q], S (&ea)[Q], S (&eb)[Q]) {
        if constexpr (ST::kDev) {
            collide2_core<float, UeqStage, FORCE, ST, OPEN>(a, UeqStage{A.u, q}, s_idx, p, rc, nr, nb, oa, ob, ea,
                                                            eb);
        } else {
            collide2_f64<FORCE>(a, A.flag, A.launch, A.watch, p, rc, nr, nb, oa, ob);
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                ea[i] = oa[i];
                eb[i] = ob[i];
            }
        }
    };

    if constexpr (two_phase<ST, OPEN, K>()) {
        // Stage s (1-based) collides row r - 3(s - 1) in iteration r and, for
        // s >= 2, gathers row r - 3(s - 1) + 1 from stage s - 1's ring.  The
        // three-row lag makes every gather read rows that earlier iterations
        // collided, so an iteration has two phases and two barriers: (A) stage
        // 1's densities of row r + 1 and the loads of row r + 2, every stage's
        // gather and densities; (B) every stage's collision.  The rings keep
        // their sizes: a ring row's last reader (phase A) and its next writer
        // (phase B) meet in the same iteration, on either side of its barrier.
        const int r_first = y0 - 2 * K + 2;                // first row stage 1 collides
        const int r1_last = y1 + 2 * K - 3;                // last row stage 1 collides
        const int r_last = y1 - 1 + 3 * (K - 1);           // the iteration of stage K's last row
        const int load_last = y1 + 2 * K - 2;              // last row loaded
        // S_n populations rotate through three arrays, the rows that stages
        // s >= 2 gather and collide through two (walk_rows)
        R pX[2][Q], pY[2][Q], pZ[2][Q];
        R sA[K - 1][2][Q], sB[K - 1][2][Q];
        // rest populations of stage s - 1 by the iteration that collided them:
        // [0] this one, [1] one earlier, [2] two earlier (the row gathered now)
        R rest[K - 1][2][3];
        // neighbourhood window: nbw[j] = bits of row r + 2 - j
        constexpr int NB = 3 * K;
        unsigned nbw[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) nbw[j] = 0u;
#pragma unroll
        for (int q = 0; q < K - 1; ++q)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                rest[q][c][0] = rest[q][c][1] = rest[q][c][2] = R(0);
#pragma unroll
                for (int i = 0; i < Q; ++i) sA[q][c][i] = sB[q][c][i] = R(0);
            }
        // prologue: densities of rows r_first - 1, r_first; populations of row
        // r_first (pX); loads of row r_first + 1 (pY) in flight
        {
            const RowAnchor an0{r_first, 0};
            const unsigned nbq = nbits(r_first - 1);
            gather1(r_first - 1, nbq, pZ);
#pragma unroll
            for (int c = 0; c < 2; ++c) L.d[0][r3(an0, r_first - 1)][c][t] = dens_w<ST>(pZ[c]);
            nbw[2] = nbits(r_first);
            gather1(r_first, nbw[2], pX);
#pragma unroll
            for (int c = 0; c < 2; ++c) L.d[0][r3(an0, r_first)][c][t] = dens_w<ST>(pX[c]);
            nbw[1] = nbits(r_first + 1);
            gather1(r_first + 1, nbw[1], pY);
            mask_words(r_first + 3);
        }
        unsigned t_r1 = tri(wr(r_first + 1), Xm, X, Xp), t_r2 = tri(wr(r_first + 2), Xm, X, Xp);

        // sg: stage s >= 2 gathers into it (the row it collides next
        // iteration); sx: the row it collides now, gathered one iteration earlier
        auto row = [&](auto ph, const int r, R (&pa)[2][Q], R (&pb)[2][Q], R (&pn)[2][Q],
                       R (&sg)[K - 1][2][Q], R (&sx)[K - 1][2][Q]) {
            const RowAnchor an{r, decltype(ph)::value};
            if constexpr (fresh_lane<ST, OPEN, K>()) asm volatile("" : "+v"(t));
            // (A) stage 1: loads of row r + 2 (past the band the last row
            // reloads load_last, cached, rather than branching round the loads);
            // densities of r + 1
            const unsigned t_r3 = tri_of_words();
            mask_words(r + 4);
            nbw[0] = neigh(t_r1, t_r2, t_r3);
            t_r1 = t_r2;
            t_r2 = t_r3;
            gather1(min(r + 2, load_last), nbw[0], pn);
#pragma unroll
            for (int c = 0; c < 2; ++c) L.d[0][r3(an, r + 1)][c][t] = dens_w<ST>(pb[c]);
            // (A) stage s >= 2: gather row g from ring q (stage s - 1), collided
            // by stage s - 1 in iterations up to r - 1
#pragma unroll
            for (int s = 2; s <= K; ++s) {
                const int q = s - 2;
                const int g = r - 3 * (s - 1) + 1;
                if (g >= y0 - 2 * (K - s) - 1 && g <= y1 + 2 * (K - s) && t >= 2 * s - 2 && t < tl - 2 * s + 2) {
                    const unsigned nbg = nbw[3 * s - 2];
                    if (OPEN || (nbg & 1u)) {
#pragma unroll
                        for (int c = 0; c < 2; ++c) {
                            sg[q][c][0] = rest[q][c][2];
#pragma unroll
                            for (int i = 1; i < Q; ++i) {
                                const bool src_fluid = OPEN || ((nbg >> opp(i)) & 1u);
                                const S v = src_fluid ? pr(an, q, g - cy(i), c, i, t - cx(i)) : pr(an, q, g, c, opp(i), t);
                                sg[q][c][i] = ST::decw(v);
                            }
                        }
                    } else {
#pragma unroll
                        for (int c = 0; c < 2; ++c)
#pragma unroll
                            for (int i = 0; i < Q; ++i) sg[q][c][i] = R(0);
                    }
#pragma unroll
                    for (int c = 0; c < 2; ++c) L.d[s - 1][r3(an, g)][c][t] = dens_w<ST>(sg[q][c]);
                }
            }
            __syncthreads();
            // (B) stage 1: collide row r into ring 0
            if (r <= r1_last && t >= 1 && t < tl - 1 && (OPEN || (nbw[2] & 1u))) {
                R nr[2][Q], rc[2], oa[Q], ob[Q];
                S ea[Q], eb[Q];
                nbr_dens(an, 0, r, nr, rc);
                collide(0, wr(r) * a.pitch + X, pa, rc, nr,
                                                      OPEN ? 0x1FFu : nbw[2], oa, ob, ea, eb);
#pragma unroll
                for (int i = 1; i < Q; ++i) {
                    pr(an, 0, r, 0, i, t) = ea[i];
                    pr(an, 0, r, 1, i, t) = eb[i];
                }
                rest[0][0][0] = ST::decw(ea[0]);
                rest[0][1][0] = ST::decw(eb[0]);
            }
            // (B) stage s >= 2: collide row h = r - 3(s - 1) from the populations
            // gathered one iteration earlier
#pragma unroll
            for (int s = 2; s <= K; ++s) {
                const int q = s - 2;
                const int h = r - 3 * (s - 1);
                const unsigned nbh = nbw[3 * s - 1];
                if (s < K) {
                    if (h >= y0 - 2 * (K - s) && h <= y1 + 2 * (K - s) - 1 && t >= 2 * s - 1 && t < tl - 2 * s + 1 &&
                        (OPEN || (nbh & 1u))) {
                        R nr[2][Q], rc[2], oa[Q], ob[Q];
                S ea[Q], eb[Q];
                        nbr_dens(an, s - 1, h, nr, rc);
                        collide(s - 1, wr(h) * a.pitch + X, sx[q], rc,
                                                              nr, OPEN ? 0x1FFu : nbh, oa, ob, ea, eb);
#pragma unroll
                        for (int i = 1; i < Q; ++i) {
                            pr(an, s - 1, h, 0, i, t) = ea[i];
                            pr(an, s - 1, h, 1, i, t) = eb[i];
                        }
                        rest[s - 1][0][0] = ST::decw(ea[0]);
                        rest[s - 1][1][0] = ST::decw(eb[0]);
                    }
                } else if (h >= y0 && h < y1 && out_lane && (OPEN || (nbh & 1u))) {
                    R nr[2][Q], rc[2], oa[Q], ob[Q];
                S ea[Q], eb[Q];
                    nbr_dens(an, s - 1, h, nr, rc);
                    collide(s - 1, h * a.pitch + X, sx[q], rc, nr,
                                                          OPEN ? 0x1FFu : nbh, oa, ob, ea, eb);
                    const OffT home = (OffT)h * pitch_b + (OffT)X * ES;
#pragma unroll
                    for (int i = 0; i < Q; ++i) {   // 32-bit offsets: a component's buffer is < 4 GiB
                        const OffT o = home + (OffT)i * plane_b;
                        ST::sce((S*)((char*)A.dst[0] + o), ea[i]);
                        ST::sce((S*)((char*)A.dst[1] + o), eb[i]);
                    }
                }
            }
            __syncthreads();
            // rotate (the populations rotate by name)
#pragma unroll
            for (int q = 0; q < K - 1; ++q)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    if constexpr (!rename_rows<ST, OPEN, K>())
#pragma unroll
                        for (int i = 0; i < Q; ++i) sx[q][c][i] = sg[q][c][i];
                    rest[q][c][2] = rest[q][c][1];
                    rest[q][c][1] = rest[q][c][0];
                }
#pragma unroll
            for (int j = NB - 1; j >= 1; --j) nbw[j] = nbw[j - 1];
        };
        // iterations r_first .. r_last
        walk_rows<rename_rows<ST, OPEN, K>()>(r_first, r_last, row, pX, pY, pZ, sA, sB);
    } else {
        // Stage s collides row r - 2(s - 1) in iteration r and, for s >= 2,
        // first gathers row r - 2s + 3: each gather reads the row stage s - 1
        // collided in the same iteration, so the stages follow each other
        // with a barrier between every gather and collision (2K - 1 per row).
        const int r_first = y0 - 2 * K + 2, r_last = y1 + 2 * K - 3;   // rows stage 1 collides
        const int load_last = y1 + 2 * K - 2;                          // last row loaded
        // S_n populations rotate through three arrays, the rows that stages
        // s >= 2 gather and collide through two (walk_rows)
        R pX[2][Q], pY[2][Q], pZ[2][Q];
        R sA[K - 1][2][Q], sB[K - 1][2][Q];
        R rest_now[K - 1][2], rest_prev[K - 1][2];
        // neighbourhood window: nbw[j] = bits of row r + 2 - j
        unsigned nbw[2 * K + 1];
#pragma unroll
        for (int j = 0; j < 2 * K + 1; ++j) nbw[j] = 0u;
#pragma unroll
        for (int q = 0; q < K - 1; ++q)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                rest_now[q][c] = rest_prev[q][c] = R(0);
#pragma unroll
                for (int i = 0; i < Q; ++i) sA[q][c][i] = sB[q][c][i] = R(0);
            }
        // prologue: densities of rows r_first - 1, r_first; populations of row
        // r_first (pX); loads of row r_first + 1 (pY) in flight
        {
            const RowAnchor an0{r_first, 0};
            const unsigned nbq = nbits(r_first - 1);
            gather1(r_first - 1, nbq, pZ);
#pragma unroll
            for (int c = 0; c < 2; ++c) L.d[0][r3(an0, r_first - 1)][c][t] = dens_w<ST>(pZ[c]);
            nbw[2] = nbits(r_first);
            gather1(r_first, nbw[2], pX);
#pragma unroll
            for (int c = 0; c < 2; ++c) L.d[0][r3(an0, r_first)][c][t] = dens_w<ST>(pX[c]);
            nbw[1] = nbits(r_first + 1);
            gather1(r_first + 1, nbw[1], pY);
            mask_words(r_first + 3);
        }
        unsigned t_r1 = tri(wr(r_first + 1), Xm, X, Xp), t_r2 = tri(wr(r_first + 2), Xm, X, Xp);

        // sg: stage s >= 2 gathers into it (the row it collides next
        // iteration); sx: the row it collides now, gathered one iteration earlier
        auto row = [&](auto ph, const int r, R (&pa)[2][Q], R (&pb)[2][Q], R (&pn)[2][Q],
                       R (&sg)[K - 1][2][Q], R (&sx)[K - 1][2][Q]) {
            const RowAnchor an{r, decltype(ph)::value};
            if constexpr (fresh_lane<ST, OPEN, K>()) asm volatile("" : "+v"(t));
            // (a) stage 1: loads of row r + 2 (the last row reloads load_last,
            // cached, rather than branching round the loads); densities of r + 1
            const unsigned t_r3 = tri_of_words();
            mask_words(r + 4);
            nbw[0] = neigh(t_r1, t_r2, t_r3);
            t_r1 = t_r2;
            t_r2 = t_r3;
            gather1(min(r + 2, load_last), nbw[0], pn);
#pragma unroll
            for (int c = 0; c < 2; ++c) L.d[0][r3(an, r + 1)][c][t] = dens_w<ST>(pb[c]);
            __syncthreads();
            // stage 1: collide row r into ring 0
            if (t >= 1 && t < tl - 1 && (OPEN || (nbw[2] & 1u))) {
                R nr[2][Q], rc[2], oa[Q], ob[Q];
                S ea[Q], eb[Q];
                nbr_dens(an, 0, r, nr, rc);
                collide(0, wr(r) * a.pitch + X, pa, rc, nr, OPEN ? 0x1FFu : nbw[2], oa, ob, ea, eb);
                if constexpr (K > 1) {
#pragma unroll
                    for (int i = 1; i < Q; ++i) {
                        pr(an, 0, r, 0, i, t) = ea[i];
                        pr(an, 0, r, 1, i, t) = eb[i];
                    }
                    rest_now[0][0] = ST::decw(ea[0]);
                    rest_now[0][1] = ST::decw(eb[0]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int s = 2; s <= K; ++s) {
                const int q = s - 2;   // register / ring index of this stage's inputs
                // gather of stage s at row g from ring q (stage s-1)
                const int g = r - 2 * s + 3;
                if (g >= y0 - 2 * (K - s) - 1 && t >= 2 * s - 2 && t < tl - 2 * s + 2) {
                    const unsigned nbg = nbw[2 * s - 1];
                    if (OPEN || (nbg & 1u)) {
#pragma unroll
                        for (int c = 0; c < 2; ++c) {
                            sg[q][c][0] = rest_prev[q][c];
#pragma unroll
                            for (int i = 1; i < Q; ++i) {
                                const bool src_fluid = OPEN || ((nbg >> opp(i)) & 1u);
                                const S v = src_fluid ? pr(an, q, g - cy(i), c, i, t - cx(i)) : pr(an, q, g, c, opp(i), t);
                                sg[q][c][i] = ST::decw(v);
                            }
                        }
                    } else {
#pragma unroll
                        for (int c = 0; c < 2; ++c)
#pragma unroll
                            for (int i = 0; i < Q; ++i) sg[q][c][i] = R(0);
                    }
#pragma unroll
                    for (int c = 0; c < 2; ++c) L.d[s - 1][r3(an, g)][c][t] = dens_w<ST>(sg[q][c]);
                }
                __syncthreads();
                // collision of stage s at row h = r - 2(s - 1)
                const int h = r - 2 * (s - 1);
                const unsigned nbh = nbw[2 * s];
                if (s < K) {
                    if (h >= y0 - 2 * (K - s) && t >= 2 * s - 1 && t < tl - 2 * s + 1 && (OPEN || (nbh & 1u))) {
                        R nr[2][Q], rc[2], oa[Q], ob[Q];
                S ea[Q], eb[Q];
                        nbr_dens(an, s - 1, h, nr, rc);
                        collide(s - 1, wr(h) * a.pitch + X, sx[q], rc, nr, OPEN ? 0x1FFu : nbh, oa, ob, ea, eb);
#pragma unroll
                        for (int i = 1; i < Q; ++i) {
                            pr(an, s - 1, h, 0, i, t) = ea[i];
                            pr(an, s - 1, h, 1, i, t) = eb[i];
                        }
                        rest_now[s - 1][0] = ST::decw(ea[0]);
                        rest_now[s - 1][1] = ST::decw(eb[0]);
                    }
                    __syncthreads();
                } else if (h >= y0 && h < y1 && out_lane && (OPEN || (nbh & 1u))) {
                    R nr[2][Q], rc[2], oa[Q], ob[Q];
                S ea[Q], eb[Q];
                    nbr_dens(an, s - 1, h, nr, rc);
                    collide(s - 1, h * a.pitch + X, sx[q], rc, nr, OPEN ? 0x1FFu : nbh, oa, ob, ea, eb);
                    const OffT home = (OffT)h * pitch_b + (OffT)X * ES;
#pragma unroll
                    for (int i = 0; i < Q; ++i) {   // 32-bit offsets: a component's buffer is < 4 GiB
                        const OffT o = home + (OffT)i * plane_b;
                        ST::sce((S*)((char*)A.dst[0] + o), ea[i]);
                        ST::sce((S*)((char*)A.dst[1] + o), eb[i]);
                    }
                }
            }
            // rotate (the populations rotate by name)
#pragma unroll
            for (int q = 0; q < K - 1; ++q)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    if constexpr (!rename_rows<ST, OPEN, K>())
#pragma unroll
                        for (int i = 0; i < Q; ++i) sx[q][c][i] = sg[q][c][i];
                    rest_prev[q][c] = rest_now[q][c];
                }
#pragma unroll
            for (int j = 2 * K; j >= 1; --j) nbw[j] = nbw[j - 1];
        };
        walk_rows<rename_rows<ST, OPEN, K>()>(r_first, r_last, row, pX, pY, pZ, sA, sB);
    }
}

// The lattice in one launch: strips of xk_out columns (A.a.gx of them) times
// bands of A.a.rows rows, workgroup ids remapped so that each XCD owns a
// contiguous range of tiles.
template <class ST, bool OPEN, int K, int T, bool FORCE>
__device__ inline void stepk_grid(const XkArgs<K, typename ST::T>& A, XkLds<ST, OPEN, K, T>& L) {
    // FP64: an earlier launch of the sequence found a site the zero-density
    // rule needs single steps for; its output is repeated by the host
    if constexpr (!ST::kDev)
        if (*G(A.flag) != 0u) return;
    const int nbx = A.a.gx;
    const int nby = (A.a.ny + A.a.rows - 1) / A.a.rows;
    int bid = blockIdx.x;
    {
        const int nblk = nbx * nby;
        const int xcd = bid & 7, k = bid >> 3;
        const int per = (nblk + 7) >> 3;
        const int full = nblk - 8 * (per - 1);
        bid = xcd < full ? xcd * per + k : full * per + (xcd - full) * (per - 1) + k;
    }
    const int bx = bid % nbx, by = bid / nbx;
    const int y0 = by * A.a.rows;
    stepk_tile<ST, OPEN, K, T, FORCE>(A, bx, y0, min(y0 + A.a.rows, A.a.ny), L);
}

template <class ST, bool OPEN, int K, int T, bool FORCE>
__global__ __launch_bounds__(T) void k_stepk(const XkArgs<K> A) {
    __shared__ XkLds<ST, OPEN, K, T> L;
    stepk_grid<ST, OPEN, K, T, FORCE>(A, L);
}

// Two steps per launch on a field: the K = 2 tile, with strips of x2_out
// columns.  (The ring of buffer sets takes 8 scalar registers where
// separate pointers for the two steps took 32, which spared the general,
// solid, variant most of its scalar spills: 9% fewer vector instructions.)
static_assert(xk_out<StoreF32, 2, kX2Threads>() == x2_out<StoreF32>() &&
                  xk_out<StoreF16, 2, kX2Threads>() == x2_out<StoreF16>(),
              "two-step strips");
template <class ST, bool OPEN, bool FORCE>
__global__ __launch_bounds__(kX2Threads) void k_step2x(const XkArgs<2> A) {
    __shared__ XkLds<ST, OPEN, 2, kX2Threads> L;
    stepk_grid<ST, OPEN, 2, kX2Threads, FORCE>(A, L);
}

// Two steps per launch in FP64 (the reference's arithmetic, bit for bit
// that of two single steps): the K = 2 tile with strips of x2d_threads
// columns, 6 fewer of them written (FP64 lines of 16 doubles split at strip
// edges merge in L2: the final stores are cached).  Without solids 192
// columns (66 KB of LDS, two workgroups per CU) held to 256 registers (two
// waves per SIMD; it fits without spilling): 4096^2 with bands of 32 rows,
// 1220 us per launch against 1305 us for one workgroup of 256 columns per
// CU.  With solids 256 columns, one wave per SIMD (at 256 registers that
// variant spills; porous 0.8: 1474 us against 1507 us).
template <bool OPEN>
__host__ __device__ constexpr int x2d_threads() {
    return OPEN ? 192 : 256;
}
template <bool OPEN>
__host__ __device__ constexpr int x2d_waves_per_eu() {
    return OPEN ? 2 : 1;
}
template <bool OPEN, bool FORCE>
__global__ __launch_bounds__(x2d_threads<OPEN>()) __attribute__((amdgpu_waves_per_eu(x2d_waves_per_eu<OPEN>()))) void
k_step2d(const XkArgs<2, double> A) {
    __shared__ XkLds<StoreF64, OPEN, 2, x2d_threads<OPEN>()> L;
    stepk_grid<StoreF64, OPEN, 2, x2d_threads<OPEN>(), FORCE>(A, L);
}

}  // namespace scl
