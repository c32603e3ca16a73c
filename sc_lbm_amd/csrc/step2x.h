// step2x.h -- strips of the two-step launches (two fused time steps per
// launch for the two-component deviation forms): threads per strip, output
// columns per strip, and the direction classes of the post-collision LDS
// ring.  The tile itself is stepk_tile with K = 2 (stepk.h): a launch reads
// P_{n-1} and writes P_{n+1}, and the intermediate populations P_n never
// leave the workgroup.  Fields launch it as k_step2x, the server's
// persistent kernel as k_serve2x.
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

}  // namespace scl
