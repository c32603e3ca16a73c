// serve.hip -- gfx950 backend of the serving layer (the scheduler is in
// host/serve_core.cpp).
//
// Memory.  One device arena of budget_bytes is allocated when the server is
// created; each resident job owns one block of it (see Layout): its bit mask,
// two population buffers per component in the server's storage form, the
// FP64 equilibrium-velocity arrays of the zero-density rule, FP64 density
// outputs, and an FP64 staging area when the job brings its own populations
// or asks for them back.
//
// Admission builds the job's streamed initial state S_0 in buffer 0 and
// scatters it into buffer 1 (P with Gather(P) = S_0), so every step of every
// job is the same gathering kernel.
//
// A round of R steps is a captured graph: R launches of k_serve over the
// work-list of all resident jobs, then the two mass reductions.  The graph
// reads everything that changes between rounds (job table, work-list, their
// sizes) through a device header that is refreshed before each launch, so a
// graph is captured once per round length and replayed for the rest of the
// server's life.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../host/serve_core.h"
#include "convert.h"
#include "geom.h"
#include "runtime.h"
#include "serve.h"
#include "step.h"

using namespace scl;
using serve::Job;
using serve::JobSpec;

namespace {

const char* hip_msg(hipError_t e) { return hipGetErrorString(e); }

constexpr int kMaxRoundSteps = 1024;   // longest round (scl_server_config.round_steps)

#define SHIP(expr)                                                                     \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            *err = std::string(#expr) + " failed: " + hip_msg(e_);                     \
            return SCL_ERR_HIP;                                                        \
        }                                                                              \
    } while (0)

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// Placement of one job inside its arena block (byte offsets).
struct Layout {
    int pitch = 0, wpr = 0;
    long long plane = 0;
    size_t mask = 0, f[2][2] = {}, ueq[kUeqRing][2][2] = {}, rho[2] = {}, stamp[2] = {}, stage = 0, total = 0;
    int nueq = 1;   // equilibrium-velocity buffer sets: 1 (FP64) or kUeqRing (deviation forms, rotating)
    bool has_stage = false;
};

Layout layout_of(const JobSpec& s, size_t es) {
    Layout L;
    L.pitch = (s.nx + 63) / 64 * 64;
    L.wpr = L.pitch / 64;
    L.plane = (long long)L.pitch * s.ny;
    size_t off = 0;
    auto take = [&](size_t n) {
        const size_t o = off;
        off += align256(n);
        return o;
    };
    L.mask = take((size_t)s.ny * L.wpr * 8);
    for (int c = 0; c < 2; ++c)
        for (int b = 0; b < 2; ++b) L.f[c][b] = take((size_t)Q * L.plane * es);
    L.nueq = es == 8 ? 1 : kUeqRing;
    for (int k = 0; k < L.nueq; ++k)
        for (int c = 0; c < 2; ++c)
            for (int d = 0; d < 2; ++d) L.ueq[k][c][d] = take((size_t)L.plane * 8);
    for (int c = 0; c < 2; ++c) L.rho[c] = take((size_t)L.plane * 8);
    // FP64: the zero-density rule's records (CompArgs::stamp), a word per
    // site, the flag and the step as of which every held value is valid
    if (es == 8)
        for (int c = 0; c < 2; ++c) L.stamp[c] = take((size_t)(L.plane + 2) * 4);
    L.has_stage = !s.f[0].empty() || (s.deliver & SCL_DELIVER_F);
    if (L.has_stage) L.stage = take((size_t)Q * L.plane * 8);
    L.total = off;
    return L;
}

// Device state of one resident job.
struct GJob {
    Layout L;
    char* base = nullptr;
    int cur = 1;   // buffer holding the post-collision populations
    Grid2 grid() const {
        return Grid2{nx, ny, L.pitch, L.wpr, L.plane, (const uint64_t*)(base + L.mask)};
    }
    void* f(int c, int b) const { return base + L.f[c][b]; }
    double* ueq(int k, int c, int d) const { return (double*)(base + L.ueq[k][c][d]); }
    int ucur = 0;   // deviation forms: buffer set holding the latest equilibrium velocities
    double* rho(int c) const { return (double*)(base + L.rho[c]); }
    unsigned* stamp(int c) const { return (unsigned*)(base + L.stamp[c]); }
    unsigned nstep = 0;   // time steps completed (numbers the next launch's steps)
    double* stage() const { return (double*)(base + L.stage); }
    int nx = 0, ny = 0;
};

dim3 site_grid(int nx, int ny) { return dim3((nx + 255) / 256, ny); }

template <class ST>
class GpuServe final : public serve::Backend {
    using T = typename ST::T;
    using S = typename ST::S;

  public:
    explicit GpuServe(const scl_server_config& cfg) : cfg_(cfg) {}
    bool device_masks() const override { return true; }
    ~GpuServe() override {
        (void)hipSetDevice(dev_);
        if (stream_) (void)hipStreamSynchronize(stream_);
        for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
        geo_.release();
        for (void* p : {(void*)arena_, (void*)hdr_d_, (void*)ctr_d_, table_d_, (void*)items_d_, (void*)partial_d_,
                        (void*)mass_d_, (void*)items2_d_, (void*)items3_d_, (void*)flags_d_, (void*)sched_d_})
            if (p) (void)hipFree(p);
        for (void* p : {(void*)hdr_h_, table_h_, (void*)items_h_, (void*)mass_h_, (void*)items2_h_, (void*)items3_h_,
                        (void*)flags_h_,
                        (void*)sched_h_})
            if (p) (void)hipHostFree(p);
        for (hipEvent_t e : ev_)
            if (e) (void)hipEventDestroy(e);
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    int init(std::string* err) {
        dev_ = runtime_device();
        SHIP(hipSetDevice(dev_));
        hipDeviceProp_t prop;
        SHIP(hipGetDeviceProperties(&prop, dev_));
        cus_ = prop.multiProcessorCount;
        SHIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));   // default priority
        SHIP(hipEventCreate(&ev_[0]));
        SHIP(hipEventCreate(&ev_[1]));
        if (cfg_.budget_bytes) {
            hipError_t e = hipMalloc(&arena_, cfg_.budget_bytes);
            if (e != hipSuccess) {
                *err = "cannot allocate a device arena of " + std::to_string(cfg_.budget_bytes) +
                       " bytes: " + hip_msg(e);
                return SCL_ERR_HIP;
            }
        }
        SHIP(hipMalloc(&hdr_d_, sizeof(ServeHdr)));
        SHIP(hipMalloc(&ctr_d_, 2 * sizeof(unsigned)));
        SHIP(hipMemsetAsync(ctr_d_, 0, 2 * sizeof(unsigned), stream_));
        SHIP(hipMalloc(&table_d_, (size_t)cfg_.max_jobs * sizeof(JobArgs<T>)));
        SHIP(hipMalloc(&mass_d_, (size_t)cfg_.max_jobs * 2 * sizeof(double)));
        SHIP(hipMalloc(&flags_d_, (size_t)cfg_.max_jobs * sizeof(unsigned)));
        SHIP(hipMalloc(&sched_d_, (size_t)3 * kMaxRoundSteps * sizeof(int)));
        SHIP(hipHostMalloc(&sched_h_, (size_t)3 * kMaxRoundSteps * sizeof(int), hipHostMallocDefault));
        SHIP(hipHostMalloc(&hdr_h_, sizeof(ServeHdr), hipHostMallocDefault));
        SHIP(hipHostMalloc(&table_h_, (size_t)cfg_.max_jobs * sizeof(JobArgs<T>), hipHostMallocDefault));
        SHIP(hipHostMalloc(&mass_h_, (size_t)cfg_.max_jobs * 2 * sizeof(double), hipHostMallocDefault));
        SHIP(hipHostMalloc(&flags_h_, (size_t)cfg_.max_jobs * sizeof(unsigned), hipHostMallocDefault));
        int rc = grow_items(4096, err);
        if (rc != SCL_OK) return rc;
        mass_grid_ = 4 * cus_;
        if (cfg_.tile_cols) {
            nw_ = cfg_.tile_cols / 64;
            rows_ = cfg_.tile_rows ? cfg_.tile_rows : 16;
            la_ = cfg_.lookahead ? cfg_.lookahead : 1;
            if ((nw_ != 2 && nw_ != 4) || nw_ * 64 != cfg_.tile_cols) {
                *err = "server tiles are 128 or 256 columns wide";
                return SCL_ERR_ARG;
            }
            if (la_ != 1 && la_ != 2) {
                *err = "lookahead is 1 or 2";
                return SCL_ERR_ARG;
            }
            if (rows_ < 1 || rows_ > 4096) {
                *err = "tile rows must be in 1..4096";
                return SCL_ERR_ARG;
            }
        } else {
            // the width is timed; a caller's rows and lookahead are kept
            if (cfg_.lookahead != 0 && cfg_.lookahead != 1 && cfg_.lookahead != 2) {
                *err = "lookahead is 1 or 2";
                return SCL_ERR_ARG;
            }
            if (cfg_.tile_rows < 0 || cfg_.tile_rows > 4096) {
                *err = "tile rows must be in 1..4096";
                return SCL_ERR_ARG;
            }
            rc = autotune(err);
            if (rc != SCL_OK) return rc;
        }
        rc = occupancy(nw_, la_, &grid_, err);
        if (rc != SCL_OK) return rc;
        if (cfg_.steps_per_launch < 0 || cfg_.steps_per_launch > 3 || cfg_.band_rows < 0 || cfg_.band_rows > 4096) {
            *err = "steps_per_launch is 0, 1, 2 or 3 and band_rows 0..4096";
            return SCL_ERR_ARG;
        }
        if constexpr (ST::kDev) {
            // multi-step launches: the deviation forms.  Automatic: three
            // per launch (porous FP32 4096^2 at porosity 0.8: 58.7k against
            // 46.5k MLUPS for two); jobs too short for three take two.
            x3_ = cfg_.steps_per_launch == 3 || cfg_.steps_per_launch == 0;
            x2_ = cfg_.steps_per_launch >= 2 || cfg_.steps_per_launch == 0;
            x2rows_ = cfg_.band_rows ? cfg_.band_rows : 32;
            x3rows_ = cfg_.band_rows ? cfg_.band_rows : 64;
            int per_cu = 0;
            SHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void*>(&k_serve2x<ST, true>), kX2Threads, 0));
            grid2_ = std::max(1, per_cu) * cus_;
            per_cu = 0;
            SHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void*>(&k_serve3x<ST, true>), x3_threads<ST, false>(), 0));
            grid3_ = std::max(1, per_cu) * cus_;
        }
        SHIP(hipStreamSynchronize(stream_));
        return SCL_OK;
    }

    int validate(const JobSpec& s, std::string* err) const override {
        const Layout L = layout_of(s, sizeof(S));
        if ((unsigned long long)Q * L.plane * sizeof(S) >= (1ull << 32)) {
            *err = "lattice too large for the server (one component's populations must stay below 4 GiB)";
            return SCL_ERR_RANGE;
        }
        return SCL_OK;
    }

    size_t footprint(const JobSpec& s) const override { return layout_of(s, sizeof(S)).total; }

    int setup(Job& j, std::string* err) override {
        SHIP(hipSetDevice(dev_));
        return build(j, arena_ + j.off, err);
    }

    int run_round(const std::vector<Job*>& jobs, const std::vector<int>& steps, int rsteps,
                  std::string* err) override {
        SHIP(hipSetDevice(dev_));
        // multi-step jobs: K steps per launch (three where enabled, else two)
        // with at least K steps in the round, kXkMinCols columns (a job
        // narrower than a strip takes one strip, whose lanes past its columns
        // and halos idle) and xk_min_rows<K>() rows
        std::vector<char> two(jobs.size(), 0), three(jobs.size(), 0);
        for (size_t k = 0; k < jobs.size(); ++k) {
            const JobSpec& sp = jobs[k]->spec;
            three[k] = x3_ && rsteps >= 3 && steps[k] >= 3 && sp.nx >= kXkMinCols && sp.ny >= xk_min_rows<3>();
            two[k] = !three[k] && x2_ && rsteps >= 2 && steps[k] >= 2 && sp.nx >= kXkMinCols &&
                     sp.ny >= xk_min_rows<2>();
        }
        // Steps a multi-step job takes singly (through k_serve): the tail
        // [tail0, steps) past its last whole group of K.
        auto tail0 = [&](int k) { return three[k] ? steps[k] / 3 * 3 : (two[k] ? steps[k] / 2 * 2 : 0); };
        auto tail = [&](int k) { return steps[k] - tail0(k); };
        // work-list order: (0) the single-step jobs, most steps first; (1) the
        // multi-step jobs with a tail, by (first tail step, tail length); (2)
        // the other multi-step jobs (their tiles only serve the mass
        // reduction).  k_serve's launch s then takes a prefix of group 0 and
        // the group-1 jobs whose tail holds s, both contiguous: tails are 1 or
        // 2 steps long, and those holding s are the ones starting at s - 1
        // with length 2 followed by those starting at s.
        std::vector<int> order(jobs.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
        auto group = [&](int k) { return !(two[k] || three[k]) ? 0 : (tail(k) ? 1 : 2); };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            if (group(a) != group(b)) return group(a) < group(b);
            if (group(a) == 0 && steps[a] != steps[b]) return steps[a] > steps[b];
            if (group(a) == 1 && tail0(a) != tail0(b)) return tail0(a) < tail0(b);
            if (group(a) == 1 && tail(a) != tail(b)) return tail(a) < tail(b);
            return jobs[a]->sites() > jobs[b]->sites();
        });
        int* sched = sched_h_;
        for (int s = 0; s < 3 * rsteps; ++s) sched[s] = 0;
        size_t need = 0, need2 = 0, need3 = 0;
        for (size_t k = 0; k < jobs.size(); ++k) {
            need += tiles_of(jobs[k]->spec.nx, jobs[k]->spec.ny, nw_, rows_);
            if (two[k]) need2 += tiles2_of(jobs[k]->spec.nx, jobs[k]->spec.ny, x2rows_);
            if (three[k]) need3 += tiles3_of(jobs[k]->spec.nx, jobs[k]->spec.ny, x3rows_);
        }
        int rc = grow_items(need, err);
        if (rc == SCL_OK) rc = grow_list(need2, &items2_d_, &items2_h_, &items2_cap_, err);
        if (rc == SCL_OK) rc = grow_list(need3, &items3_d_, &items3_h_, &items3_cap_, err);
        if (rc != SCL_OK) return rc;
        JobArgs<T>* table = (JobArgs<T>*)table_h_;
        int n = 0, n2 = 0, n3 = 0;
        bool force2 = false, force3 = false;   // a two- / three-step job of the round has a body force
        for (int k : order) {
            const GJob* g = (const GJob*)jobs[k]->dev;
            JobArgs<T>& ja = table[k];
            ja.a = args_of(jobs[k]->spec, *g);
            ja.cur = g->cur;
            ja.steps = steps[k];
            ja.x2 = two[k];
            ja.x3 = three[k];
            ja.fin = g->cur ^ (launches(steps[k], per_launch(two[k], three[k])) & 1);
            fill_ueq(ja, *g);
            ja.item_begin = n;
            n = emit_tiles(k, g->nx, g->ny, nw_, rows_, items_h_, n);
            ja.item_end = n;
            if (group(k) == 0) {
                for (int s = 0; s < steps[k]; ++s) sched[3 * s] = n;
            } else if (group(k) == 1) {
                for (int s = tail0(k); s < steps[k]; ++s) {
                    if (sched[3 * s + 2] == 0) sched[3 * s + 1] = ja.item_begin;
                    sched[3 * s + 2] = n;
                }
            }
            const bool force = jobs[k]->spec.has_force[0] || jobs[k]->spec.has_force[1];
            if (two[k]) {
                n2 = emit_tiles_k(k, g->nx, g->ny, x2rows_, kOut2, items2_h_, n2);
                force2 = force2 || force;
            }
            if (three[k]) {
                n3 = emit_tiles_k(k, g->nx, g->ny, x3rows_, kOut3, items3_h_, n3);
                force3 = force3 || force;
            }
        }
        *hdr_h_ = ServeHdr{table_d_, items_d_, partial_d_, mass_d_, n, (int)jobs.size(), 64 * nw_,
                           items2_d_, n2, items3_d_, n3, flags_d_, sched_d_};
        SHIP(hipMemcpyAsync(sched_d_, sched_h_, (size_t)3 * rsteps * sizeof(int), hipMemcpyHostToDevice, stream_));
        SHIP(hipMemcpyAsync(table_d_, table_h_, jobs.size() * sizeof(JobArgs<T>), hipMemcpyHostToDevice, stream_));
        SHIP(hipMemcpyAsync(items_d_, items_h_, (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
        if (n2) SHIP(hipMemcpyAsync(items2_d_, items2_h_, (size_t)n2 * sizeof(uint64_t), hipMemcpyHostToDevice,
                                    stream_));
        if (n3) SHIP(hipMemcpyAsync(items3_d_, items3_h_, (size_t)n3 * sizeof(uint64_t), hipMemcpyHostToDevice,
                                    stream_));
        SHIP(hipMemcpyAsync(hdr_d_, hdr_h_, sizeof(ServeHdr), hipMemcpyHostToDevice, stream_));
        hipGraphExec_t ge = nullptr;
        const Multi mk{n2 > 0, force2, n3 > 0, force3};
        if (cfg_.use_graphs) {
            rc = graph(rsteps, mk, &ge, err);
            if (rc != SCL_OK) return rc;
        }
        SHIP(hipEventRecord(ev_[0], stream_));
        if (ge) {
            SHIP(hipGraphLaunch(ge, stream_));
        } else {
            launch_round(rsteps, mk);
            SHIP(hipGetLastError());
        }
        SHIP(hipEventRecord(ev_[1], stream_));
        SHIP(hipMemcpyAsync(mass_h_, mass_d_, jobs.size() * 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
        SHIP(hipMemcpyAsync(flags_h_, flags_d_, jobs.size() * sizeof(unsigned), hipMemcpyDeviceToHost, stream_));
        SHIP(hipStreamSynchronize(stream_));
        float ms = 0;
        SHIP(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
        device_ms_.store(device_ms_.load() + ms);
        for (size_t k = 0; k < jobs.size(); ++k) {
            GJob* g = (GJob*)jobs[k]->dev;
            jobs[k]->mass[0] = mass_h_[2 * k];
            jobs[k]->mass[1] = mass_h_[2 * k + 1];
            g->cur ^= launches(steps[k], per_launch(two[k], three[k])) & 1;
            g->ucur = (g->ucur + steps[k]) % g->L.nueq;
            g->nstep += (unsigned)steps[k];
            if (two[k] || three[k]) pair_updates_ += (long long)tail0(k) * g->nx * g->ny;
            if (flags_h_[k])
                jobs[k]->device_error = "zero-density rule: a site's density fell from >= 1e-8 to <= 1e-16 "
                                        "within one step without its previous equilibrium velocity held";
        }
        return SCL_OK;
    }

    int finish(Job& j, std::string* err) override {
        SHIP(hipSetDevice(dev_));
        const GJob* g = (const GJob*)j.dev;
        const Grid2 gr = g->grid();
        const size_t n = (size_t)g->nx * g->ny;
        for (int c = 0; c < 2; ++c) {
            const double rho0 = ST::kDev ? j.spec.rho0[c] : 0.0;
            const S* P = (const S*)g->f(c, g->cur);
            if (j.spec.deliver & SCL_DELIVER_RHO) {
                hipLaunchKernelGGL(k_job_density<ST>, site_grid(g->nx, g->ny), dim3(256), 0, stream_, gr, P, rho0,
                                   g->rho(c));
                SHIP(hipGetLastError());
                j.rho[c].resize(n);
                SHIP(hipMemcpy2DAsync(j.rho[c].data(), g->nx * 8, g->rho(c), (size_t)g->L.pitch * 8, g->nx * 8,
                                      g->ny, hipMemcpyDeviceToHost, stream_));
            }
            if (j.spec.deliver & SCL_DELIVER_F) {
                hipLaunchKernelGGL((k_to_f64<ST, true>), site_grid(g->nx, g->ny), dim3(256), 0, stream_, gr, P, rho0,
                                   g->stage());
                SHIP(hipGetLastError());
                j.f[c].resize(9 * n);
                SHIP(hipMemcpy2DAsync(j.f[c].data(), g->nx * 8, g->stage(), (size_t)g->L.pitch * 8, g->nx * 8,
                                      (size_t)9 * g->ny, hipMemcpyDeviceToHost, stream_));
            }
            SHIP(hipStreamSynchronize(stream_));
        }
        return SCL_OK;
    }

    void drop(Job& j) override {
        delete (GJob*)j.dev;
        j.dev = nullptr;
    }

    void shape(int* cols, int* rows, int* la) const override {
        *cols = 64 * nw_;
        *rows = rows_;
        *la = la_;
    }
    int workgroups() const override { return grid_; }
    int steps_per_launch() const override { return x3_ ? 3 : (x2_ ? 2 : 1); }
    double device_ms() const override { return device_ms_.load(); }
    long long pair_updates() const override { return pair_updates_.load(); }

  private:
    static void fill_ueq(JobArgs<T>& ja, const GJob& g) {
        for (int k = 0; k < kUeqRing; ++k)
            for (int c = 0; c < 2; ++c)
                for (int d = 0; d < 2; ++d) ja.ueq[k][c][d] = g.ueq(k < g.L.nueq ? k : 0, c, d);
        ja.ucur = g.ucur;
        // The ring view (two-step tiles).  layout_of takes the buffers in
        // [k][c][d] order, each plane * 8 bytes, a multiple of 512 (the pitch
        // is a multiple of 64), so they are contiguous: component c's set k,
        // direction d sits at ueq(0, c, 0) + (4k + d) * plane doubles.
        UeqRing& r = ja.ring;
        r.n = g.L.plane;
        r.set_stride = 4 * g.L.plane;
        r.nsets = g.L.nueq;
        for (int c = 0; c < 2; ++c) {
            r.base[c] = g.ueq(0, c, 0);
            r.cur[c] = g.ucur;
        }
    }

    static size_t tiles_of(int nx, int ny, int nw, int rows) {
        return (size_t)((nx + 64 * nw - 1) / (64 * nw)) * (size_t)((ny + rows - 1) / rows);
    }

    // Steps per launch of a job in a round, and the launches that swap its
    // buffers in a round of `steps` steps: steps / K groups and the rest
    // one at a time.
    static int per_launch(bool two, bool three) { return three ? 3 : (two ? 2 : 1); }
    static int launches(int steps, int k) { return steps / k + steps % k; }

    // Narrowest job of the multi-step kernels: a strip's halo lanes wrap at
    // most once (2 K - 1 <= 5 columns each side).
    static constexpr int kXkMinCols = 8;
    // output columns of a two- / three-step strip in this storage form (the tiles')
    static constexpr int kOut2 = x2_out<ST>();
    static constexpr int kOut3 = xk_out<ST, 3, x3_threads<ST, false>()>();
    static size_t tiles2_of(int nx, int ny, int rows) {
        return (size_t)((nx + kOut2 - 1) / kOut2) * (size_t)((ny + rows - 1) / rows);
    }
    static size_t tiles3_of(int nx, int ny, int rows) {
        return (size_t)((nx + kOut3 - 1) / kOut3) * (size_t)((ny + rows - 1) / rows);
    }
    static int emit_tiles_k(int k, int nx, int ny, int rows, int out_cols, uint64_t* out, int n) {
        const int gx = (nx + out_cols - 1) / out_cols;
        for (int y0 = 0; y0 < ny; y0 += rows)
            for (int bx = 0; bx < gx; ++bx) out[n++] = pack_item(k, bx, y0, std::min(y0 + rows, ny));
        return n;
    }

    // A multi-step tile list (device and pinned host copies) of at least
    // `need` entries.
    int grow_list(size_t need, uint64_t** dev, uint64_t** host, size_t* cap_io, std::string* err) {
        if (need <= *cap_io) return SCL_OK;
        const size_t cap = std::max<size_t>(*cap_io * 2, need);
        if (*dev) SHIP(hipFree(*dev));
        if (*host) SHIP(hipHostFree(*host));
        *dev = nullptr;
        *host = nullptr;
        SHIP(hipMalloc(dev, cap * sizeof(uint64_t)));
        SHIP(hipHostMalloc(host, cap * sizeof(uint64_t), hipHostMallocDefault));
        *cap_io = cap;
        return SCL_OK;
    }

    // Tiles of job k, band by band, strips left to right within a band.
    static int emit_tiles(int k, int nx, int ny, int nw, int rows, uint64_t* out, int n) {
        const int gx = (nx + 64 * nw - 1) / (64 * nw);
        for (int y0 = 0; y0 < ny; y0 += rows)
            for (int bx = 0; bx < gx; ++bx) out[n++] = pack_item(k, bx, y0, std::min(y0 + rows, ny));
        return n;
    }

    int grow_items(size_t need, std::string* err) {
        if (need <= items_cap_) return SCL_OK;
        size_t cap = std::max<size_t>(items_cap_ * 2, need);
        if (items_d_) SHIP(hipFree(items_d_));
        if (partial_d_) SHIP(hipFree(partial_d_));
        if (items_h_) SHIP(hipHostFree(items_h_));
        items_d_ = nullptr;
        partial_d_ = nullptr;
        items_h_ = nullptr;
        SHIP(hipMalloc(&items_d_, cap * sizeof(uint64_t)));
        SHIP(hipMalloc(&partial_d_, cap * 2 * sizeof(double)));
        SHIP(hipHostMalloc(&items_h_, cap * sizeof(uint64_t), hipHostMallocDefault));
        items_cap_ = cap;
        return SCL_OK;
    }

    // Per-component parameters exactly as the single-lattice path forms them
    // (omega = 1/tau, include/fluid.h:62; 1 - omega and 1/omega,
    // src/lbm.cpp:213-226, 284, 308; -1.0 * G, src/lbm.cpp:163-170 and
    // src/fluid.cpp:154-160).
    StepArgs<T> args_of(const JobSpec& s, const GJob& g) const {
        StepArgs<T> a{};
        a.nx = g.nx;
        a.ny = g.ny;
        a.pitch = g.L.pitch;
        a.wpr = g.L.wpr;
        a.plane = g.L.plane;
        a.mask = (const uint64_t*)(g.base + g.L.mask);
        a.rows = rows_;
        a.gx = (g.nx + 64 * nw_ - 1) / (64 * nw_);
        for (int c = 0; c < 2; ++c) {
            CompArgs<T>& ca = a.c[c];
            const double omega = 1 / s.tau[c];
            ca.src = g.f(c, 0);
            ca.dst = g.f(c, 1);
            ca.omega = (T)omega;
            ca.omc = (T)(1.0 - omega);
            ca.inv_omega = (T)(1.0 / omega);
            ca.gneg = (T)(-1.0 * s.g_fluid[c]);
            ca.gsneg = (T)(-1.0 * s.g_solid[c]);
            ca.rho0 = (T)(ST::kDev ? s.rho0[c] : 0.0);
            set_force(ca, s.has_force[c] ? s.force[c] : nullptr, omega);
            ca.step = g.nstep + 1;
            ca.stamp = ST::kDev ? nullptr : g.stamp(c);
            ca.ueqx = g.ueq(0, c, 0);
            ca.ueqy = g.ueq(0, c, 1);
            ca.ueqx_in = ca.ueqx;
            ca.ueqy_in = ca.ueqy;
        }
        return a;
    }

    // The job's lattice at `base`: zeroed block, mask, S_0 in buffer 0 and
    // its scatter in buffer 1.
    int build(Job& j, char* base, std::string* err) {
        const JobSpec& s = j.spec;
        auto g = std::unique_ptr<GJob>(new GJob);
        g->L = layout_of(s, sizeof(S));
        g->base = base;
        g->nx = s.nx;
        g->ny = s.ny;
        const Layout& L = g->L;
        if (!j.prepared) {   // jobs built outside the scheduler (the autotuner's)
            if (s.mask.empty()) {
                j.fluid_sites = serve::job_disks(s, &j.disks);
                j.device_mask = true;
            } else {
                serve::pack_mask(s.mask, s.nx, s.ny, &j.bits);
            }
            j.prepared = true;
        }
        SHIP(hipMemsetAsync(base, 0, L.total, stream_));
        if (j.device_mask) {
            const int rc = raster_job(j, (uint64_t*)(base + L.mask), err);
            if (rc != SCL_OK) return rc;
        } else {
            if ((int)(j.bits.size() / s.ny) != L.wpr || j.bits.size() != (size_t)s.ny * L.wpr) {
                *err = "packed mask does not match the lattice";
                return SCL_ERR_ARG;
            }
            SHIP(hipMemcpyAsync(base + L.mask, j.bits.data(), j.bits.size() * 8, hipMemcpyHostToDevice, stream_));
        }
        if (!j.saved.empty()) {   // a resumed job: its saved state (save) into buffer 0 and set 0
            if (j.saved.size() != saved_bytes(L)) {
                *err = "saved state does not match the lattice";
                return SCL_ERR_ARG;
            }
            const char* in = j.saved.data();
            const size_t pop = (size_t)Q * L.plane * sizeof(S), ueq = (size_t)L.plane * 8;
            for (int c = 0; c < 2; ++c, in += pop)
                SHIP(hipMemcpyAsync(g->f(c, 0), in, pop, hipMemcpyHostToDevice, stream_));
            for (int c = 0; c < 2; ++c)
                for (int d = 0; d < 2; ++d, in += ueq)
                    SHIP(hipMemcpyAsync(g->ueq(0, c, d), in, ueq, hipMemcpyHostToDevice, stream_));
            // the steps continue their numbering; every held equilibrium
            // velocity restored is valid as of the last step done
            g->nstep = (unsigned)j.steps_done;
            const unsigned valid = g->nstep;
            if (!ST::kDev)
                for (int c = 0; c < 2; ++c)
                    SHIP(hipMemcpyAsync(g->stamp(c) + L.plane + 1, &valid, sizeof valid, hipMemcpyHostToDevice,
                                        stream_));
            SHIP(hipStreamSynchronize(stream_));
            j.saved.clear();
            g->cur = 0;
            g->ucur = 0;
            j.dev = g.release();
            return SCL_OK;
        }
        const Grid2 gr = g->grid();
        const dim3 sg = site_grid(s.nx, s.ny);
        for (int c = 0; c < 2; ++c) {
            const double rho0 = ST::kDev ? s.rho0[c] : 0.0;
            S* s0 = (S*)g->f(c, 0);
            if (!s.f[0].empty()) {
                SHIP(hipMemcpy2DAsync(g->stage(), (size_t)L.pitch * 8, s.f[c].data(), (size_t)s.nx * 8,
                                      (size_t)s.nx * 8, (size_t)9 * s.ny, hipMemcpyHostToDevice, stream_));
                hipLaunchKernelGGL(k_from_f64<ST>, sg, dim3(256), 0, stream_, gr, g->stage(), rho0, s0);
            } else {
                const serve::IcParams p{s.nx, s.ny, s.ic, s.rho_hi, s.rho_lo, s.radius, s.noise, s.ic_seed};
                hipLaunchKernelGGL(k_init_ic<ST>, sg, dim3(256), 0, stream_, gr, p, c, rho0, s0);
            }
            SHIP(hipGetLastError());
            hipLaunchKernelGGL(k_scatter<ST>, sg, dim3(256), 0, stream_, gr, (const S*)s0, (S*)g->f(c, 1));
            SHIP(hipGetLastError());
        }
        SHIP(hipStreamSynchronize(stream_));   // host vectors above must outlive the copies
        g->cur = 1;
        j.dev = g.release();
        return SCL_OK;
    }

    // A generated mask on the device: the sampler's disks (walls: rows 0 and
    // ny - 1) rasterised into `bits`, its fluid sites counted and checked
    // against the sampler's count.
    int raster_job(Job& j, uint64_t* bits, std::string* err) {
        const JobSpec& s = j.spec;
        static_assert(sizeof(serve::Disk) == sizeof(RasterDisk), "the sampler's disks are the raster's circles");
        SHIP(geo_.upload(reinterpret_cast<const RasterDisk*>(j.disks.data()), j.disks.size(), stream_));
        SHIP(raster_mask(stream_, bits, s.nx, s.ny, s.walls ? 1 : 0, 1, geo_.disks(), geo_.rows(),
                         (int)j.disks.size(), geo_.n_rows, geo_.count()));
        unsigned long long n = 0;
        SHIP(hipMemcpyAsync(&n, geo_.count(), sizeof n, hipMemcpyDeviceToHost, stream_));
        SHIP(hipStreamSynchronize(stream_));
        if ((long long)n != j.fluid_sites) {
            *err = "device mask has " + std::to_string(n) + " fluid sites, the sampler counted " +
                   std::to_string(j.fluid_sites);
            return SCL_ERR_HIP;
        }
        // The disks stay with the job (a few KB): a suspended job's lattice
        // is rebuilt from them when it resumes.
        return SCL_OK;
    }

    // Pre-emption: the populations of buffer cur (post-collision, as stored)
    // and the equilibrium-velocity set the next step reads, per component.
    static size_t saved_bytes(const Layout& L) {
        return 2 * (size_t)Q * L.plane * sizeof(S) + 4 * (size_t)L.plane * 8;
    }
    bool can_save() const override { return true; }
    int save(Job& j, std::vector<char>* out, std::string* err) override {
        SHIP(hipSetDevice(dev_));
        const GJob* g = (const GJob*)j.dev;
        const Layout& L = g->L;
        const size_t pop = (size_t)Q * L.plane * sizeof(S), ueq = (size_t)L.plane * 8;
        const int k = L.nueq > 1 ? g->ucur : 0;
        out->resize(saved_bytes(L));
        char* o = out->data();
        for (int c = 0; c < 2; ++c, o += pop)
            SHIP(hipMemcpyAsync(o, g->f(c, g->cur), pop, hipMemcpyDeviceToHost, stream_));
        for (int c = 0; c < 2; ++c)
            for (int d = 0; d < 2; ++d, o += ueq)
                SHIP(hipMemcpyAsync(o, g->ueq(k, c, d), ueq, hipMemcpyDeviceToHost, stream_));
        SHIP(hipStreamSynchronize(stream_));
        return SCL_OK;
    }

    template <int NW, int LA>
    void launch_one(int grid, int s) {
        hipLaunchKernelGGL((k_serve<ST, NW, LA>), dim3(grid), dim3(64 * NW), 0, stream_, hdr_d_, s, ctr_d_);
    }

    void launch_step(int nw, int la, int grid, int s) {
        if (nw == 2) {
            if (la == 2) launch_one<2, 2>(grid, s);
            else launch_one<2, 1>(grid, s);
        } else {
            if (la == 2) launch_one<4, 2>(grid, s);
            else launch_one<4, 1>(grid, s);
        }
    }

    static const void* kernel_of(int nw, int la) {
        if (nw == 2)
            return la == 2 ? reinterpret_cast<const void*>(&k_serve<ST, 2, 2>)
                           : reinterpret_cast<const void*>(&k_serve<ST, 2, 1>);
        return la == 2 ? reinterpret_cast<const void*>(&k_serve<ST, 4, 2>)
                       : reinterpret_cast<const void*>(&k_serve<ST, 4, 1>);
    }

    // Workgroups of the persistent grid: as many as are resident at once.
    int occupancy(int nw, int la, int* grid, std::string* err) {
        int per_cu = 0;
        SHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_of(nw, la), 64 * nw, 0));
        *grid = std::max(1, per_cu) * cus_;
        return SCL_OK;
    }

    // The multi-step launches of a round: two- / three-step jobs present, and
    // whether any of them has a body force.
    struct Multi {
        bool two, force2, three, force3;
    };

    // The round as a graph: rsteps single-step launches, with (two) per pair
    // q a two-step launch before the single-step launch of step 2q and
    // (three) per triple q a three-step launch before that of step 3q; the
    // force flags pick the body-force variants.
    int graph(int rsteps, const Multi& mk, hipGraphExec_t* out, std::string* err) {
        const int key = 16 * rsteps + (mk.two ? 1 : 0) + (mk.two && mk.force2 ? 2 : 0) + (mk.three ? 4 : 0) +
                        (mk.three && mk.force3 ? 8 : 0);
        auto it = graphs_.find(key);
        if (it != graphs_.end()) {
            *out = it->second;
            return SCL_OK;
        }
        SHIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        launch_round(rsteps, mk);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(stream_, &gr);
        if (e != hipSuccess) {
            *err = std::string("graph capture failed: ") + hip_msg(e);
            return SCL_ERR_HIP;
        }
        hipGraphExec_t ge = nullptr;
        e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) {
            *err = std::string("graph instantiation failed: ") + hip_msg(e);
            return SCL_ERR_HIP;
        }
        graphs_[key] = ge;
        *out = ge;
        return SCL_OK;
    }

    // The launches of a round: `rsteps` steps (pairs of steps through
    // k_serve2x, triples through k_serve3x where such jobs are present),
    // then the per-job FP64 mass.
    void launch_round(int rsteps, const Multi& mk) {
        for (int s = 0; s < rsteps; ++s) {
            if (mk.three && s % 3 == 0 && s + 2 < rsteps) launch_triple(s / 3, mk.force3);
            if (mk.two && !(s & 1) && s + 1 < rsteps) launch_pair(s / 2, mk.force2);
            launch_step(nw_, la_, grid_, s);
        }
        hipLaunchKernelGGL(k_serve_mass<ST>, dim3(mass_grid_), dim3(256), 0, stream_, hdr_d_);
        hipLaunchKernelGGL(k_serve_mass_final<T>, dim3(cfg_.max_jobs), dim3(256), 0, stream_, hdr_d_);
    }

    void launch_triple(int q, bool force) {
        if constexpr (ST::kDev) {
            constexpr int T3 = x3_threads<ST, false>();
            if (force)
                hipLaunchKernelGGL((k_serve3x<ST, true>), dim3(grid3_), dim3(T3), 0, stream_, hdr_d_, q, ctr_d_);
            else
                hipLaunchKernelGGL((k_serve3x<ST, false>), dim3(grid3_), dim3(T3), 0, stream_, hdr_d_, q, ctr_d_);
        }
    }

    void launch_pair(int q, bool force) {
        if constexpr (ST::kDev) {
            if (force)
                hipLaunchKernelGGL((k_serve2x<ST, true>), dim3(grid2_), dim3(kX2Threads), 0, stream_, hdr_d_, q,
                                   ctr_d_);
            else
                hipLaunchKernelGGL((k_serve2x<ST, false>), dim3(grid2_), dim3(kX2Threads), 0, stream_, hdr_d_, q,
                                   ctr_d_);
        }
    }

    // Tile shape by measurement: two resident 2048 x 2048 porous jobs
    // (porosity 0.8, disks of radius 8..20, droplet initial state) in
    // scratch memory; every (width 128/256, lookahead 1/2) of 16-row tiles
    // runs 2 + 8 steps and the fastest mean step wins.  The rows are not
    // tuned on the probe: taller tiles win on two large jobs but leave the
    // small lattices of a served mix short of workgroups.  On a mix of
    // sides 128..2048 the capacity was 34.4k-35.1k / 33.0k-33.4k / 30.2k
    // MLUPS with 16 / 32 / 64 rows (256 columns, lookahead 2), and the
    // probe had chosen 64, then 32 once 64 was excluded.
    int autotune(std::string* err) {
        JobSpec s;
        s.nx = s.ny = 2048;
        s.porosity = 0.8;
        s.rmin = 8;
        s.rmax = 20;
        s.geom_seed = 99;
        s.ic = SCL_IC_DROPLET;
        s.rho_hi = 1.0;
        s.rho_lo = 0.06;
        s.radius = 400;
        s.noise = 0.01;
        s.ic_seed = 7;
        s.tau[0] = s.tau[1] = 1.0;
        s.g_fluid[0] = s.g_fluid[1] = 0.9;
        const size_t bytes = layout_of(s, sizeof(S)).total;
        char* scratch = nullptr;
        hipError_t e = hipMalloc(&scratch, 2 * bytes);
        if (e != hipSuccess) {
            *err = std::string("autotune scratch allocation failed: ") + hip_msg(e);
            return SCL_ERR_HIP;
        }
        Job ja, jb;
        ja.spec = s;
        jb.spec = s;
        ja.spec.ic_seed = 11;
        int rc = build(ja, scratch, err);
        if (rc == SCL_OK) rc = build(jb, scratch + bytes, err);
        std::vector<Job*> jobs = {&ja, &jb};
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (rc == SCL_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
            *err = "event creation failed";
            rc = SCL_ERR_HIP;
        }
        float best = 1e30f;
        // rows: the caller's, else 16; lookahead: the caller's, else timed
        const int rows_fixed = cfg_.tile_rows ? cfg_.tile_rows : 16;
        int bnw = 4, brows = rows_fixed, bla = cfg_.lookahead ? cfg_.lookahead : 1;
        std::vector<int> las = {1, 2};
        if (cfg_.lookahead) las = {cfg_.lookahead};
        const int nws[] = {2, 4}, rowss[] = {rows_fixed};
        for (int nw : nws)
            for (int la : las)
                for (int rows : rowss) {
                    if (rc != SCL_OK) break;
                    int grid = 0;
                    rc = occupancy(nw, la, &grid, err);
                    if (rc != SCL_OK) break;
                    nw_ = nw;
                    rows_ = rows;
                    la_ = la;
                    const std::vector<int> steps = {10, 10};
                    size_t need = 2 * tiles_of(s.nx, s.ny, nw, rows);
                    rc = grow_items(need, err);
                    if (rc != SCL_OK) break;
                    JobArgs<T>* table = (JobArgs<T>*)table_h_;
                    int n = 0;
                    for (int k = 0; k < 2; ++k) {
                        const GJob* g = (const GJob*)jobs[k]->dev;
                        table[k].a = args_of(jobs[k]->spec, *g);
                        table[k].cur = g->cur;
                        table[k].steps = 10;
                        table[k].x2 = 0;
                        table[k].x3 = 0;
                        table[k].fin = g->cur;
                        fill_ueq(table[k], *g);
                        table[k].item_begin = n;
                        n = emit_tiles(k, g->nx, g->ny, nw, rows, items_h_, n);
                        table[k].item_end = n;
                    }
                    *hdr_h_ = ServeHdr{table_d_, items_d_, partial_d_, mass_d_, n, 2, 64 * nw, items2_d_, 0,
                                       items3_d_, 0, flags_d_, sched_d_};
                    for (int s = 0; s < 10; ++s) {
                        sched_h_[3 * s] = n;
                        sched_h_[3 * s + 1] = sched_h_[3 * s + 2] = 0;
                    }
                    if (hipMemcpyAsync(table_d_, table_h_, 2 * sizeof(JobArgs<T>), hipMemcpyHostToDevice,
                                       stream_) != hipSuccess ||
                        hipMemcpyAsync(items_d_, items_h_, (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice,
                                       stream_) != hipSuccess ||
                        hipMemcpyAsync(hdr_d_, hdr_h_, sizeof(ServeHdr), hipMemcpyHostToDevice, stream_) !=
                            hipSuccess ||
                        hipMemcpyAsync(sched_d_, sched_h_, 30 * sizeof(int), hipMemcpyHostToDevice, stream_) !=
                            hipSuccess) {
                        *err = "autotune upload failed";
                        rc = SCL_ERR_HIP;
                        break;
                    }
                    launch_step(nw, la, grid, 0);
                    launch_step(nw, la, grid, 1);
                    (void)hipEventRecord(e0, stream_);
                    for (int t = 2; t < 10; ++t) launch_step(nw, la, grid, t);
                    (void)hipEventRecord(e1, stream_);
                    if (hipStreamSynchronize(stream_) != hipSuccess || hipGetLastError() != hipSuccess) {
                        *err = "autotune launch failed";
                        rc = SCL_ERR_HIP;
                        break;
                    }
                    float ms = 0;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    // the two scratch jobs took 10 steps: both buffers swap back
                    if (ms < best) {
                        best = ms;
                        bnw = nw;
                        brows = rows;
                        bla = la;
                    }
                }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        drop(ja);
        drop(jb);
        (void)hipStreamSynchronize(stream_);
        (void)hipFree(scratch);
        nw_ = bnw;
        rows_ = brows;
        la_ = bla;
        return rc;
    }

    scl_server_config cfg_;
    int dev_ = 0, cus_ = 256;
    hipStream_t stream_ = nullptr;
    char* arena_ = nullptr;
    ServeHdr* hdr_d_ = nullptr;
    ServeHdr* hdr_h_ = nullptr;
    int* sched_d_ = nullptr;   // k_serve's item ranges per step of the round (ServeHdr::sched)
    int* sched_h_ = nullptr;
    GeoStage geo_;             // disks of the mask being rasterised
    unsigned* ctr_d_ = nullptr;
    void* table_d_ = nullptr;
    void* table_h_ = nullptr;
    uint64_t* items_d_ = nullptr;
    uint64_t* items_h_ = nullptr;
    size_t items_cap_ = 0;
    double* partial_d_ = nullptr;
    double* mass_d_ = nullptr;
    double* mass_h_ = nullptr;
    unsigned* flags_d_ = nullptr;   // zero-density rule flag per job of the round
    unsigned* flags_h_ = nullptr;
    int nw_ = 4, rows_ = 16, la_ = 1, grid_ = 0, mass_grid_ = 1024;
    bool x2_ = false;                  // two steps per launch (deviation forms)
    int x2rows_ = 32, grid2_ = 0;      // band rows of the two-step tiles; its persistent grid
    uint64_t* items2_d_ = nullptr;
    uint64_t* items2_h_ = nullptr;
    size_t items2_cap_ = 0;
    bool x3_ = false;                  // three steps per launch (deviation forms)
    int x3rows_ = 64, grid3_ = 0;      // band rows of the three-step tiles; its persistent grid
    uint64_t* items3_d_ = nullptr;
    uint64_t* items3_h_ = nullptr;
    size_t items3_cap_ = 0;
    std::map<int, hipGraphExec_t> graphs_;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    std::atomic<double> device_ms_{0.0};   // written by the request loop, read by stats
    std::atomic<long long> pair_updates_{0};
};

template <class ST>
int make_backend(const scl_server_config& cfg, std::unique_ptr<serve::Backend>* out, std::string* err) {
    auto b = std::unique_ptr<GpuServe<ST>>(new GpuServe<ST>(cfg));
    const int rc = b->init(err);
    if (rc != SCL_OK) return rc;
    *out = std::move(b);
    return SCL_OK;
}

}  // namespace

extern "C" int scl_server_create(const scl_server_config* cfg, scl_server** out) {
    if (!out) return serve::report(SCL_ERR_ARG, "null output");
    *out = nullptr;
    if (!cfg) return serve::report(SCL_ERR_ARG, "null config");
    const scl_server_config& c = *cfg;
    if (c.precision < SCL_FP64 || c.precision > SCL_FP16_DEV)
        return serve::report(SCL_ERR_ARG, "bad precision " + std::to_string(c.precision));
    if (c.max_jobs < 1 || c.max_jobs > 65535) return serve::report(SCL_ERR_ARG, "max_jobs must be in 1..65535");
    if (c.max_pending < 1) return serve::report(SCL_ERR_ARG, "max_pending must be >= 1");
    if (c.round_steps < 1 || c.round_steps > kMaxRoundSteps)
        return serve::report(SCL_ERR_ARG, "round_steps must be in 1..1024");
    if (!(c.mass_tol >= 0)) return serve::report(SCL_ERR_ARG, "mass_tol must be >= 0 (0: by precision)");
    if (c.max_bypass < 0) return serve::report(SCL_ERR_ARG, "max_bypass must be >= 0");
    int rc = runtime_ready();
    if (rc != SCL_OK) return rc;
    std::unique_ptr<serve::Backend> be;
    std::string err;
    switch (c.precision) {
        case SCL_FP64: rc = make_backend<StoreF64>(c, &be, &err); break;
        case SCL_FP32_DEV: rc = make_backend<StoreF32>(c, &be, &err); break;
        default: rc = make_backend<StoreF16>(c, &be, &err); break;
    }
    if (rc != SCL_OK) return serve::report(rc, err);
    auto s = std::unique_ptr<scl_server>(new scl_server);
    s->sched.reset(new serve::Scheduler(std::move(be), c));
    *out = s.release();
    return SCL_OK;
}

// ---------------------------------------------------------------------------
// The geometry pipeline on the device, as calls of its own
// ---------------------------------------------------------------------------
namespace {

// Rasterises walls and circles into an ny x ceil(nx/64) bit mask on the
// library's device, copies it to `bits` and counts its fluid sites.
int raster_to_host(int nx, int ny, int dh, int where_x, const std::vector<RasterDisk>& disks, uint64_t* bits,
                   size_t n_words, long long* fluid, double* dev_ms, std::string* err) {
    const int wpr = (nx + 63) / 64;
    if (n_words != (size_t)ny * wpr) {
        *err = "bits must hold ny * ceil(nx / 64) = " + std::to_string((size_t)ny * wpr) + " words";
        return SCL_ERR_ARG;
    }
    int rc = runtime_ready();
    if (rc != SCL_OK) return rc;
    SHIP(hipSetDevice(runtime_device()));
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    uint64_t* bits_d = nullptr;
    GeoStage geo;
    unsigned long long n = 0;
    float ms = 0.f;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);   // default priority
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMalloc(&bits_d, n_words * sizeof(uint64_t));
    if (e == hipSuccess) e = geo.upload(disks.data(), disks.size(), st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    if (e == hipSuccess)
        e = raster_mask(st, bits_d, nx, ny, dh, where_x, geo.disks(), geo.rows(), (int)disks.size(), geo.n_rows,
                        geo.count());
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipMemcpyAsync(bits, bits_d, n_words * sizeof(uint64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&n, geo.count(), sizeof n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (st) (void)hipStreamSynchronize(st);
    geo.release();
    if (bits_d) (void)hipFree(bits_d);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (e != hipSuccess) {
        *err = std::string("device geometry failed: ") + hip_msg(e);
        return SCL_ERR_HIP;
    }
    if (fluid) *fluid = (long long)n;
    if (dev_ms) *dev_ms = ms;
    return SCL_OK;
}

}  // namespace

extern "C" int scl_mask_raster(int nx, int ny, long long wall_thickness, int walls_along, const long long* circles,
                               int n_circles, uint64_t* bits, size_t n_words, long long* fluid_sites,
                               double* device_ms) {
    if (nx < 1 || ny < 1 || n_circles < 0 || (n_circles > 0 && !circles) || !bits)
        return serve::report(SCL_ERR_ARG, "bad lattice size, circle list or output");
    if (walls_along < 0 || walls_along > 2) return serve::report(SCL_ERR_ARG, "walls_along is 0, 1 (x) or 2 (y)");
    if (wall_thickness < 0) return serve::report(SCL_ERR_ARG, "negative wall thickness");
    // Geometry::add_walls (src/geometry.cpp:34-67)
    if (walls_along == 2 && wall_thickness > nx)
        return serve::report(SCL_ERR_ARG, "Wall thickness larger than domain width (x)");
    if (walls_along == 1 && wall_thickness > ny)
        return serve::report(SCL_ERR_ARG, "Wall thickness larger than domain height (y)");
    std::vector<RasterDisk> disks((size_t)n_circles);
    for (int k = 0; k < n_circles; ++k) {
        const long long D = circles[3 * k], xc = circles[3 * k + 1], yc = circles[3 * k + 2];
        if (D < 1 || xc < 0 || yc < 0) return serve::report(SCL_ERR_ARG, "circle " + std::to_string(k) + ": bad size or centre");
        // Geometry::check_object_bounds (src/geometry.cpp:150-162)
        if (xc - D / 2 < 0 || yc - D / 2 < 0)
            return serve::report(SCL_ERR_ARG, "Circle lower bounds do not fit the domain.");
        if (xc + D / 2 >= nx || yc + D / 2 >= ny)
            return serve::report(SCL_ERR_ARG, "Circle upper bounds do not fit the domain.");
        const long long L = D % 2 == 0 ? D - 1 : D;
        if (L > kMaxRasterL) return serve::report(SCL_ERR_RANGE, "circle diameter above 65535");
        disks[(size_t)k] = RasterDisk{(int)xc, (int)yc, (int)L};
    }
    std::string err;
    const int rc = raster_to_host(nx, ny, walls_along ? (int)wall_thickness : 0, walls_along == 1, disks, bits, n_words,
                                  fluid_sites, device_ms, &err);
    return rc == SCL_OK ? rc : serve::report(rc, err);
}

extern "C" int scl_job_mask_bits(const scl_job* job, uint64_t* bits, size_t n_words, long long* fluid_sites,
                                 double* ms) {
    if (!job || !bits) return serve::report(SCL_ERR_ARG, "null job or output");
    JobSpec spec;
    std::string err;
    int rc = serve::spec_from_job(job, &spec, &err);
    if (rc != SCL_OK) return serve::report(rc, err);
    const int wpr = (spec.nx + 63) / 64;
    if (n_words != (size_t)spec.ny * wpr) return serve::report(SCL_ERR_ARG, "bits must hold ny * ceil(nx / 64) words");
    if (!spec.mask.empty()) {   // the caller's mask, packed
        std::vector<uint64_t> b;
        serve::pack_mask(spec.mask, spec.nx, spec.ny, &b);
        std::memcpy(bits, b.data(), n_words * sizeof(uint64_t));
        long long nf = 0;
        for (uint64_t w : b) nf += __builtin_popcountll(w);
        if (fluid_sites) *fluid_sites = nf;
        if (ms) ms[0] = ms[1] = 0.0;
        return SCL_OK;
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<serve::Disk> d;
    const long long expect = serve::job_disks(spec, &d);
    const double sample_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::vector<RasterDisk> disks(d.size());
    for (size_t k = 0; k < d.size(); ++k) disks[k] = RasterDisk{d[k].xc, d[k].yc, d[k].L};
    long long nf = 0;
    double dev = 0.0;
    rc = raster_to_host(spec.nx, spec.ny, spec.walls ? 1 : 0, 1, disks, bits, n_words, &nf, &dev, &err);
    if (rc != SCL_OK) return serve::report(rc, err);
    if (nf != expect)
        return serve::report(SCL_ERR_HIP, "device mask has " + std::to_string(nf) + " fluid sites, the sampler counted " +
                                              std::to_string(expect));
    if (fluid_sites) *fluid_sites = nf;
    if (ms) {
        ms[0] = sample_ms;
        ms[1] = dev;
    }
    return SCL_OK;
}
