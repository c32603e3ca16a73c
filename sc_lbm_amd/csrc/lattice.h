// lattice.h -- D2Q9 constants, device layout and population storage forms
// shared by the gfx950 kernels.
//
// Device layout (all arrays resident in HBM):
//   pitch   = nx rounded up to a multiple of 64 sites; site (x, y) lives at
//             y*pitch + x in every per-site array.
//   plane   = elements between consecutive population directions; population
//             i of site s is at i*plane + s (structure of arrays, so a
//             wavefront touches 64 consecutive sites of one direction).
//   mask    = one bit per site, 1 = fluid, packed little-endian in 64-bit
//             words, pitch/64 words per row.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scl {

constexpr int Q = 9;
// Lattice velocities, direction order rest, E, N, W, S, NE, NW, SW, SE
// (reference include/lbm.h:126-128).
__host__ __device__ constexpr int cx(int i) { return i == 1 || i == 5 || i == 8 ? 1 : (i == 3 || i == 6 || i == 7 ? -1 : 0); }
__host__ __device__ constexpr int cy(int i) { return i == 2 || i == 5 || i == 6 ? 1 : (i == 4 || i == 7 || i == 8 ? -1 : 0); }
// Opposite direction (reference bounce-back rules, include/lbm.h:124).
__host__ __device__ constexpr int opp(int i) { return i == 0 ? 0 : (i <= 4 ? ((i + 1) % 4) + 1 : ((i - 3) % 4) + 5); }
// Interaction weights for the fluid-fluid and fluid-solid sums
// (include/lbm.h:118-122): 1/9 axis, 1/36 diagonal.
__host__ __device__ constexpr double wint(int i) { return i == 0 ? 0.0 : (i <= 4 ? 1.0 / 9 : 1.0 / 36); }
// Equilibrium weights (include/fluid.h:282).
__host__ __device__ constexpr double weq(int i) { return i == 0 ? 4.0 / 9.0 : (i <= 4 ? 1.0 / 9.0 : 1.0 / 36.0); }

static_assert(opp(1) == 3 && opp(2) == 4 && opp(3) == 1 && opp(4) == 2, "opp axis");
static_assert(opp(5) == 7 && opp(6) == 8 && opp(7) == 5 && opp(8) == 6, "opp diag");

// Density below which the reference keeps the previous equilibrium velocity
// (include/utils.h:145-150 with tol 1e-16, src/lbm.cpp:262-269).
constexpr double kZeroTol = 1e-16;
// Fused kernels hold u_eq in HBM only where a component's density is below
// a threshold (the only sites where the next step can fall back to it):
// 1e-8 in FP64.  The deviation forms hold it below a threshold above the
// noise their storage leaves in the density of an empty component
// (StoreF32 / StoreF16 ::kUeqWatch: 1e-5 and 2e-3, against about 1e-7 and
// 5e-4 at rho0 = 1), where the density can fall to 0 and back in one step.
// Held sites cost a store and an atomic per step, so the thresholds stay
// below the densities of the dilute phases of the configurations (a 5e-2
// threshold halved the FP16 throughput of the 4096^2 spinodal lattice).
constexpr double kUeqWatch64 = 1e-8;
constexpr double kUeqWatch = kUeqWatch64;

// FP16 deviation scale: stored value = (f_i - w_i*rho0) * 2^8.
constexpr float kHalfScale = 256.0f;
constexpr float kHalfInv = 1.0f / 256.0f;

// Global-memory view of a pointer.  A pointer that reaches a kernel through
// memory (a job table) rather than as a kernel argument has no known address
// space, and the compiler would access it with flat instructions, which also
// wait on LDS traffic; every population, mask and per-site array access goes
// through this cast so it is a global one.
// (The qualifiers apply to the device compilation; the host pass only
// parses device code.)
#if defined(__HIP_DEVICE_COMPILE__)
#define SCL_AS_GLOBAL __attribute__((address_space(1)))
#define SCL_AS_CONST __attribute__((address_space(4)))
#else
#define SCL_AS_GLOBAL
#define SCL_AS_CONST
#endif
template <class T>
__device__ inline SCL_AS_GLOBAL T* G(T* p) {
    return (SCL_AS_GLOBAL T*)p;
}
template <class T>
__device__ inline const SCL_AS_GLOBAL T* G(const T* p) {
    return (const SCL_AS_GLOBAL T*)p;
}

// Constant-memory view of a table that no thread writes during the kernel
// (uniform reads of it become scalar loads).
template <class T>
__device__ inline const SCL_AS_CONST T* K(const T* p) {
    return (const SCL_AS_CONST T*)p;
}

// ---------------------------------------------------------------------------
// Storage forms.  T is the arithmetic type, S the stored type.  The
// deviation forms store d_i = f_i - w_i*rho0.
// ---------------------------------------------------------------------------
// Population stores are streamed (non-temporal): a step's output is read
// back only by the next step, after the whole lattice has passed through the
// caches.
#ifndef SCL_STREAM_STORES
#define SCL_STREAM_STORES 1
#endif
template <class S>
__device__ inline void stream_store(S* p, S v) {
#if SCL_STREAM_STORES
    __builtin_nontemporal_store(v, G(p));
#else
    *G(p) = v;
#endif
}

// st: streamed store; stc: cached store (for scattered partial-line writes).
// Working units (the fused kernels' arithmetic on the populations): the
// stored value decoded without its scale, i.e. the deviation times kUnit, a
// power of two, so every product and sum in those units is the one on the
// deviations scaled exactly (ldw / decw / encw; densities, velocities and
// forces stay in plain units).
struct StoreF64 {
    using S = double;
    using T = double;
    static constexpr bool kDev = false;
    static constexpr T kUnit = 1.0;
    __device__ static inline T ldw(const S* p) { return *G(p); }
    __device__ static inline T decw(S s) { return s; }
    __device__ static inline S encw(T v) { return v; }
    static constexpr T kUeqWatch = kUeqWatch64;
    __device__ static inline T ld(const S* p) { return *G(p); }
    __device__ static inline void st(S* p, T v) { stream_store<S>(p, v); }
    __device__ static inline void stc(S* p, T v) { *G(p) = v; }
    __device__ static inline void se(S* p, S e) { stream_store<S>(p, e); }   // store an encoded value
    __device__ static inline void sce(S* p, S e) { *G(p) = e; }
    __device__ static inline S enc(T v) { return v; }   // value as stored
    __device__ static inline T dec(S s) { return s; }   // stored value as loaded
    __device__ static inline T rnd(T v) { return v; }   // v rounded to what the store keeps
};

struct StoreF32 {
    using S = float;
    using T = float;
    static constexpr bool kDev = true;
    static constexpr T kUnit = 1.0f;
    __device__ static inline T ldw(const S* p) { return *G(p); }
    __device__ static inline T decw(S s) { return s; }
    __device__ static inline S encw(T v) { return v; }
    static constexpr T kUeqWatch = 1e-5f;
    __device__ static inline T ld(const S* p) { return *G(p); }
    __device__ static inline void st(S* p, T v) { stream_store<S>(p, v); }
    __device__ static inline void stc(S* p, T v) { *G(p) = v; }
    __device__ static inline void se(S* p, S e) { stream_store<S>(p, e); }   // store an encoded value
    __device__ static inline void sce(S* p, S e) { *G(p) = e; }
    __device__ static inline S enc(T v) { return v; }
    __device__ static inline T dec(S s) { return s; }
    __device__ static inline T rnd(T v) { return v; }
};

struct StoreF16 {
    using S = _Float16;
    using T = float;
    static constexpr bool kDev = true;
    static constexpr T kUnit = kHalfScale;   // the stored value itself: one conversion each way
    __device__ static inline T ldw(const S* p) { return (float)(*G(p)); }
    __device__ static inline T decw(S s) { return (float)s; }
    __device__ static inline S encw(T v) { return (_Float16)v; }
    static constexpr T kUeqWatch = 2e-3f;
    __device__ static inline T ld(const S* p) { return (float)(*G(p)) * kHalfInv; }
    __device__ static inline void st(S* p, T v) { stream_store<S>(p, (_Float16)(v * kHalfScale)); }
    __device__ static inline void stc(S* p, T v) { *G(p) = (_Float16)(v * kHalfScale); }
    __device__ static inline void se(S* p, S e) { stream_store<S>(p, e); }   // store an encoded value
    __device__ static inline void sce(S* p, S e) { *G(p) = e; }
    __device__ static inline S enc(T v) { return (_Float16)(v * kHalfScale); }
    __device__ static inline T dec(S s) { return (float)s * kHalfInv; }
    __device__ static inline T rnd(T v) { return dec(enc(v)); }
};

// Wrap a column / row index that is at most one step outside [0, n).
__device__ inline int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ inline int fluid_bit(const uint64_t* __restrict__ mask, int wpr, int row, int col) {
    return (int)((G(mask)[(size_t)row * wpr + (col >> 6)] >> (col & 63)) & 1ull);
}

// Bits (col-1, col, col+1) of one row: bit0 west, bit1 centre, bit2 east.
__device__ inline unsigned triple(const uint64_t* __restrict__ mask, int wpr, int row, int colm,
                                  int col, int colp) {
    return (unsigned)fluid_bit(mask, wpr, row, colm) | ((unsigned)fluid_bit(mask, wpr, row, col) << 1) |
           ((unsigned)fluid_bit(mask, wpr, row, colp) << 2);
}

// 9-bit neighbourhood: bit i = fluid(x + c_i, y + c_i) from the triples of
// rows y-1 (s), y (c) and y+1 (n).
__device__ inline unsigned neigh_bits(unsigned s, unsigned c, unsigned n) {
    return ((c >> 1) & 1u)               // 0 rest
           | (((c >> 2) & 1u) << 1)      // 1 E
           | (((n >> 1) & 1u) << 2)      // 2 N
           | ((c & 1u) << 3)             // 3 W
           | (((s >> 1) & 1u) << 4)      // 4 S
           | (((n >> 2) & 1u) << 5)      // 5 NE
           | ((n & 1u) << 6)             // 6 NW
           | ((s & 1u) << 7)             // 7 SW
           | (((s >> 2) & 1u) << 8);     // 8 SE
}

}  // namespace scl
