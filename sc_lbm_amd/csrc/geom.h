// geom.h -- the geometry pipeline on the device: an all-fluid bit mask with
// add_walls' solid rows or columns, add_circle node sets cleared from it, and
// its fluid sites counted.
//
// Bit mask layout (the kernels' own): row-major, wpr = ceil(nx / 64) 64-bit
// words per row, bit x & 63 of word x >> 6 set where site x is fluid; bits
// past nx stay zero.
//
// Node sets (src/geom_object/ellipse.cpp:12-33 as add_circle calls it,
// src/geometry.cpp:132-147): a circle of odd diameter L (an even one has
// already shrunk by one) centred on (xc, yc) holds node (x, y) of the box
// xc +- L/2, yc +- L/2 where
//     (dx^2 L^2) / 4 + (dy^2 L^2) / 4 <= L^4 / 16
// in unsigned 64-bit arithmetic, each division truncating.  For L < 2^16 no
// product wraps, and the left side grows with |dx|: the nodes of row y are
// the run |dx| <= m(dy), which k_mask_disks finds by bisection and clears
// with 64-bit atomic ANDs (circles may share words, and may overlap).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace scl {

constexpr int kMaxRasterL = 65535;   // circles up to this diameter (no wrapped products)

struct RasterDisk {
    int xc, yc, L;   // centre, odd diameter
};

// Every word of the lattice: fluid bits of the columns < nx, minus the walls
// (where_x: rows y < dh and y >= ny - dh; else columns x < dh and x >= nx - dh).
static __global__ void k_mask_base(uint64_t* __restrict__ bits, int nx, int ny, int wpr, int dh, int where_x) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)ny * wpr) return;
    const int y = (int)(i / wpr), w = (int)(i % wpr);
    const int c0 = 64 * w;
    auto run = [](int lo, int hi) -> uint64_t {   // bits lo..hi-1 of a word, 0 <= lo <= hi <= 64
        if (hi <= lo) return 0ull;
        const uint64_t top = hi == 64 ? ~0ull : ((1ull << hi) - 1ull);
        return top & ~((1ull << lo) - 1ull);
    };
    uint64_t word = run(0, min(64, nx - c0));
    if (dh > 0) {
        if (where_x) {
            if (y < dh || y >= ny - dh) word = 0;
        } else {
            word &= ~run(0, max(0, min(64, dh - c0)));
            word &= ~run(max(0, min(64, nx - dh - c0)), 64);
        }
    }
    bits[i] = word;
}

// Left side of the node test for offsets (dx, dy) of a circle of diameter L.
__device__ inline uint64_t circle_lhs(uint64_t dx, uint64_t dy, uint64_t L) {
    return dx * dx * L * L / 4 + dy * dy * L * L / 4;
}

// One thread per (circle, row of its box): clears the row's nodes.  rows[k]
// is the first row index of circle k (prefix sums of the diameters), n_rows
// their total.
static __global__ void k_mask_disks(uint64_t* __restrict__ bits, int wpr, const RasterDisk* __restrict__ disks,
                                    const long long* __restrict__ rows, int n_disks, long long n_rows) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_rows) return;
    int lo = 0, hi = n_disks - 1;   // the circle whose rows hold i: last k with rows[k] <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rows[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const RasterDisk d = disks[lo];
    const int h = d.L / 2;
    const int y = d.yc - h + (int)(i - rows[lo]);
    const uint64_t L = (uint64_t)d.L, rhs = L * L * L * L / 16;
    const uint64_t ady = (uint64_t)(y >= d.yc ? y - d.yc : d.yc - y);
    if (circle_lhs(0, ady, L) > rhs) return;   // no node in this row
    int a = 0, b = h;                          // largest m <= h with lhs(m, dy) <= rhs
    while (a < b) {
        const int m = (a + b + 1) >> 1;
        if (circle_lhs((uint64_t)m, ady, L) <= rhs) a = m;
        else b = m - 1;
    }
    const int x0 = d.xc - a, x1 = d.xc + a;    // inclusive run
    uint64_t* row = bits + (size_t)y * wpr;
    for (int w = x0 >> 6; w <= (x1 >> 6); ++w) {
        const int l = max(x0 - 64 * w, 0), r = min(x1 - 64 * w, 63);
        const uint64_t top = r == 63 ? ~0ull : ((1ull << (r + 1)) - 1ull);
        const uint64_t clear = top & ~((1ull << l) - 1ull);
        atomicAnd((unsigned long long*)(row + w), (unsigned long long)~clear);
    }
}

// Fluid sites: popcount of every word into *count (zero beforehand).
static __global__ void k_mask_count(const uint64_t* __restrict__ bits, long long n_words,
                                    unsigned long long* __restrict__ count) {
    __shared__ unsigned long long part;
    if (threadIdx.x == 0) part = 0;
    __syncthreads();
    unsigned long long c = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_words;
         i += (long long)gridDim.x * blockDim.x)
        c += (unsigned long long)__popcll(bits[i]);
    atomicAdd(&part, c);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(count, part);
}

// The three launches on `stream`: base words, circles (disks / rows already
// on the device), count into *count_d (cleared here).
inline hipError_t raster_mask(hipStream_t stream, uint64_t* bits_d, int nx, int ny, int dh, int where_x,
                              const RasterDisk* disks_d, const long long* rows_d, int n_disks, long long n_rows,
                              unsigned long long* count_d) {
    const int wpr = (nx + 63) / 64;
    const long long words = (long long)ny * wpr;
    hipLaunchKernelGGL(k_mask_base, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, stream, bits_d, nx, ny, wpr,
                       dh, where_x);
    if (n_disks > 0)
        hipLaunchKernelGGL(k_mask_disks, dim3((unsigned)((n_rows + 255) / 256)), dim3(256), 0, stream, bits_d, wpr,
                           disks_d, rows_d, n_disks, n_rows);
    hipError_t e = hipMemsetAsync(count_d, 0, sizeof(unsigned long long), stream);
    if (e != hipSuccess) return e;
    const long long blocks = std::min<long long>((words + 255) / 256, 2048);
    hipLaunchKernelGGL(k_mask_count, dim3((unsigned)std::max<long long>(blocks, 1)), dim3(256), 0, stream, bits_d,
                       words, count_d);
    return hipGetLastError();
}

// Device staging of a circle list: the circles, their first rows and the
// fluid count, in one allocation grown as needed.  The host copies that
// upload() makes stay until the next upload; callers synchronise the stream
// before uploading again.
struct GeoStage {
    char* d = nullptr;
    size_t cap = 0, n = 0;
    long long n_rows = 0;
    std::vector<RasterDisk> h_disks;
    std::vector<long long> h_rows;

    RasterDisk* disks() const { return (RasterDisk*)(d + 256); }
    long long* rows() const { return (long long*)(d + 256 + align(n * sizeof(RasterDisk))); }
    unsigned long long* count() const { return (unsigned long long*)d; }
    static size_t align(size_t b) { return (b + 255) / 256 * 256; }

    hipError_t upload(const RasterDisk* src, size_t count_, hipStream_t stream) {
        n = count_;
        h_disks.assign(src, src + n);
        h_rows.resize(n);
        n_rows = 0;
        for (size_t k = 0; k < n; ++k) {
            h_rows[k] = n_rows;
            n_rows += h_disks[k].L;
        }
        const size_t need = 256 + align(n * sizeof(RasterDisk)) + align(n * sizeof(long long));
        if (need > cap) {
            const size_t c = std::max(need, 2 * cap);
            if (d) {
                hipError_t e = hipStreamSynchronize(stream);
                if (e == hipSuccess) e = hipFree(d);
                if (e != hipSuccess) return e;
                d = nullptr;
                cap = 0;
            }
            hipError_t e = hipMalloc(&d, c);
            if (e != hipSuccess) return e;
            cap = c;
        }
        if (n == 0) return hipSuccess;
        hipError_t e = hipMemcpyAsync(disks(), h_disks.data(), n * sizeof(RasterDisk), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(rows(), h_rows.data(), n * sizeof(long long), hipMemcpyHostToDevice, stream);
        return e;
    }
    void release() {
        if (d) (void)hipFree(d);
        d = nullptr;
        cap = 0;
    }
};

}  // namespace scl
