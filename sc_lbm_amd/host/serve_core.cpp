// serve_core.cpp -- scheduler of the serving layer (see serve_core.h) and
// the backend-independent part of its C-ABI (include/sc_lbm_amd.h).
//
// Request loop.  One thread owns the backend.  Each pass:
//   1. drops queued jobs that were cancelled or whose time-out passed, and
//      marks resident ones for eviction;
//   2. admits released jobs in (release, id) order while a slot is free and
//      the arena can hold the job (a job whose mask is still being built
//      holds the later ones back, so the order does not depend on thread
//      timing).  A job that does not fit is passed over,
//      until `max_bypass` later jobs have been admitted ahead of it; then
//      admission waits for it (no starvation of large jobs);
//   3. builds the new jobs' lattices, then runs one round: every resident
//      job advances min(remaining, R) steps, R = the round length rounded to
//      a power of two no larger than round_steps;
//   4. checks each job's FP64 mass (finite, relative drift from the first
//      round within mass_tol), finishes completed jobs (deliverables) and
//      releases their memory.
// Submissions, status queries and waits take the lock only briefly; backend
// work runs outside it.
#include "serve_core.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "serve_ic.h"

namespace scl {
namespace serve {

// ---------------------------------------------------------------------------
// Arena
// ---------------------------------------------------------------------------
Arena::Arena(size_t size, size_t align) : size_(size / align * align), align_(align) {
    if (size_) free_[0] = size_;
}

bool Arena::alloc(size_t bytes, size_t* off) {
    const size_t n = rounded(bytes == 0 ? 1 : bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < n) continue;
        const size_t o = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[o + n] = len - n;
        live_[o] = n;
        used_ += n;
        peak_ = std::max(peak_, used_);
        *off = o;
        return true;
    }
    return false;
}

void Arena::release(size_t off) {
    auto it = live_.find(off);
    if (it == live_.end()) return;
    size_t o = off, n = it->second;
    live_.erase(it);
    used_ -= n;
    auto nx = free_.lower_bound(o);
    if (nx != free_.end() && o + n == nx->first) {   // merge with the following block
        n += nx->second;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {                        // merge with the preceding block
        auto pv = std::prev(nx);
        if (pv->first + pv->second == o) {
            pv->second += n;
            return;
        }
    }
    free_[o] = n;
}

size_t Arena::largest_free() const {
    size_t m = 0;
    for (const auto& kv : free_) m = std::max(m, kv.second);
    return m;
}

// ---------------------------------------------------------------------------
// Job specification
// ---------------------------------------------------------------------------
int spec_from_job(const scl_job* j, JobSpec* s, std::string* err) {
    if (!j) {
        *err = "null job";
        return SCL_ERR_ARG;
    }
    if (j->nx < 1 || j->ny < 1 || j->nx > 32767 || j->ny > 32767) {
        *err = "lattice " + std::to_string(j->nx) + "x" + std::to_string(j->ny) + " outside 1..32767";
        return SCL_ERR_ARG;
    }
    if (j->steps < 1) {
        *err = "steps must be >= 1";
        return SCL_ERR_ARG;
    }
    for (int c = 0; c < 2; ++c) {
        if (!(j->tau[c] > 0.0) || !std::isfinite(j->tau[c])) {
            *err = "tau must be positive";
            return SCL_ERR_ARG;
        }
    }
    if (!(j->release_ms >= 0.0) || !(j->timeout_ms >= 0.0)) {
        *err = "release_ms and timeout_ms must be >= 0";
        return SCL_ERR_ARG;
    }
    if (j->deliver & ~(SCL_DELIVER_RHO | SCL_DELIVER_F)) {
        *err = "unknown deliver bits";
        return SCL_ERR_ARG;
    }
    const size_t n = (size_t)j->nx * j->ny;
    s->nx = j->nx;
    s->ny = j->ny;
    if (j->mask) {
        s->mask.assign(j->mask, j->mask + n);
        for (int& v : s->mask) v = v ? 1 : 0;
    } else {
        if (!(j->porosity > 0.0 && j->porosity <= 1.0)) {
            *err = "porosity must be in (0, 1]";
            return SCL_ERR_ARG;
        }
        if (j->porosity < 1.0 && (j->rmin < 0 || j->rmax < j->rmin)) {
            *err = "disk radii need 0 <= rmin <= rmax";
            return SCL_ERR_ARG;
        }
        s->porosity = j->porosity;
        s->rmin = j->rmin;
        s->rmax = j->rmax;
        s->walls = j->walls ? 1 : 0;
        s->geom_seed = j->geom_seed;
        if (s->walls && j->ny < 3) {
            *err = "walls need ny >= 3";
            return SCL_ERR_ARG;
        }
    }
    if ((j->f[0] == nullptr) != (j->f[1] == nullptr)) {
        *err = "give initial populations for both components or for none";
        return SCL_ERR_ARG;
    }
    if (j->f[0]) {
        for (int c = 0; c < 2; ++c) s->f[c].assign(j->f[c], j->f[c] + 9 * n);
    } else {
        if (j->ic < SCL_IC_UNIFORM || j->ic > SCL_IC_STRATIFIED) {
            *err = "unknown initial condition";
            return SCL_ERR_ARG;
        }
        s->ic = j->ic;
        s->rho_hi = j->rho_hi;
        s->rho_lo = j->rho_lo;
        s->radius = j->radius;
        s->noise = j->noise;
        s->ic_seed = j->ic_seed;
    }
    for (int c = 0; c < 2; ++c) {
        s->tau[c] = j->tau[c];
        s->g_fluid[c] = j->g_fluid[c];
        s->g_solid[c] = j->g_solid[c];
        s->has_force[c] = j->force[c] != nullptr;
        for (int i = 0; i < 9; ++i) s->force[c][i] = j->force[c] ? j->force[c][i] : 0.0;
    }
    s->steps = j->steps;
    s->release_ms = j->release_ms;
    s->timeout_ms = j->timeout_ms;
    s->deliver = j->deliver;
    // Deviation reference density: storage spacing scales with |f - w rho0|,
    // so the middle of the densities present keeps the worst phase finest.
    for (int c = 0; c < 2; ++c) {
        double r = j->rho0[c];
        if (!(r >= 0.0) || !std::isfinite(r)) {
            *err = "rho0 must be finite and >= 0";
            return SCL_ERR_ARG;
        }
        if (r == 0.0) {
            if (j->f[0]) {
                double sum = 0.0;
                long long cnt = 0;
                for (size_t k = 0; k < n; ++k) {
                    double d = 0.0;
                    for (int i = 0; i < 9; ++i) d += s->f[c][(size_t)i * n + k];
                    if (d != 0.0) {
                        sum += d;
                        ++cnt;
                    }
                }
                r = cnt ? sum / cnt : 1.0;
            } else {
                r = 0.5 * (s->rho_hi + s->rho_lo);
            }
        }
        r = std::round(r * 1024.0) / 1024.0;
        s->rho0[c] = r > 0.0 && r < 1e6 ? r : 1.0;
    }
    return SCL_OK;
}

namespace {

// Sequential uniform draws: blocks of 4096 SplitMix64 outputs, block k
// seeded with seed + k * 0x632BE59BD9B4E019.
class Draws {
  public:
    explicit Draws(uint64_t seed) : seed_(seed) {}
    double next() {
        if (pos_ == 4096) {
            ++blk_;
            pos_ = 0;
        }
        return uniform01_at(seed_ + blk_ * 0x632BE59BD9B4E019ull, pos_++);
    }

  private:
    uint64_t seed_, blk_ = 0, pos_ = 0;
};

// add_circle node set (src/geometry.cpp:132-147 with
// src/geom_object/ellipse.cpp:12-33) of odd diameter L: node (ix, iy) of the
// box xc +- L/2, yc +- L/2 is solid when
// ((ix-xc)^2*L^2)/4 + ((iy-yc)^2*L^2)/4 <= L^4/16 in integer arithmetic.
void carve_disk(std::vector<int>& m, int nx, const Disk& d) {
    const long long L = d.L, h = L / 2, L2 = L * L, rhs = L2 * L2 / 16;
    for (long long y = d.yc - h; y <= d.yc + h; ++y)
        for (long long x = d.xc - h; x <= d.xc + h; ++x) {
            const long long lhs = ((x - d.xc) * (x - d.xc) * L2) / 4 + ((y - d.yc) * (y - d.yc) * L2) / 4;
            if (lhs <= rhs) m[(size_t)y * nx + (size_t)x] = 0;
        }
}

}  // namespace

long long disk_nodes(long long L) {
    const long long h = L / 2, L2 = L * L, rhs = L2 * L2 / 16;
    long long n = 0;
    for (long long dy = -h; dy <= h; ++dy)
        for (long long dx = -h; dx <= h; ++dx) n += (dx * dx * L2) / 4 + (dy * dy * L2) / 4 <= rhs;
    return n;
}

void job_mask(const JobSpec& s, std::vector<int>* out) {
    const int nx = s.nx, ny = s.ny;
    if (!s.mask.empty()) {
        *out = s.mask;
        return;
    }
    std::vector<Disk> disks;
    job_disks(s, &disks);
    std::vector<int>& m = *out;
    m.assign((size_t)nx * ny, 1);
    if (s.walls)
        for (int x = 0; x < nx; ++x) {
            m[x] = 0;
            m[(size_t)(ny - 1) * nx + x] = 0;
        }
    for (const Disk& d : disks) carve_disk(m, nx, d);
}

long long job_disks(const JobSpec& s, std::vector<Disk>* out) {
    const int nx = s.nx, ny = s.ny;
    out->clear();
    long long fluid = (long long)nx * ny;
    if (s.walls) fluid -= 2LL * nx;
    if (s.porosity >= 1.0) return fluid;
    const double total = (double)nx * (double)ny;
    const int y_lo = s.walls ? 1 : 0, y_hi = s.walls ? ny - 2 : ny - 1;
    // disks in a uniform grid of cells at least as wide as the largest
    // centre distance that can overlap (r + r2 + 1 <= 2 rmax + 1)
    const int cell = 2 * s.rmax + 2;
    const int gx = nx / cell + 1, gy = ny / cell + 1;
    std::vector<std::vector<int>> grid((size_t)gx * gy);
    struct Placed {
        long long r, xc, yc;
    };
    std::vector<Placed> disks;
    std::vector<long long> nodes(s.rmax + 1, -1);   // node count by radius
    Draws st(s.geom_seed);
    const int max_tries = 200000;
    for (int tries = 0; (double)fluid / total > s.porosity && tries < max_tries;) {
        ++tries;
        long long r = s.rmin + (long long)(st.next() * (double)(s.rmax - s.rmin + 1));
        r = std::min<long long>(r, s.rmax);
        const long long lo_x = r, hi_x = nx - 1 - r, lo_y = y_lo + r, hi_y = y_hi - r;
        if (hi_x < lo_x || hi_y < lo_y) continue;
        const long long xc = lo_x + (long long)(st.next() * (double)(hi_x - lo_x + 1));
        const long long yc = lo_y + (long long)(st.next() * (double)(hi_y - lo_y + 1));
        const int cx = (int)(xc / cell), cy = (int)(yc / cell);
        bool ok = true;
        for (int j = std::max(0, cy - 1); j <= std::min(gy - 1, cy + 1) && ok; ++j)
            for (int i = std::max(0, cx - 1); i <= std::min(gx - 1, cx + 1) && ok; ++i)
                for (int k : grid[(size_t)j * gx + i]) {
                    const Placed& o = disks[k];
                    const long long dx = xc - o.xc, dy = yc - o.yc, rr = r + o.r + 1;
                    if (dx * dx + dy * dy <= rr * rr) {
                        ok = false;
                        break;
                    }
                }
        if (!ok) continue;
        if (nodes[r] < 0) nodes[r] = disk_nodes(2 * r + 1);
        fluid -= nodes[r];
        grid[(size_t)cy * gx + cx].push_back((int)disks.size());
        disks.push_back(Placed{r, xc, yc});
        out->push_back(Disk{(int)xc, (int)yc, (int)(2 * r + 1)});
    }
    return fluid;
}

void pack_mask(const std::vector<int>& m, int nx, int ny, std::vector<uint64_t>* bits) {
    const int wpr = (nx + 63) / 64;
    bits->assign((size_t)ny * wpr, 0);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
            if (m[(size_t)y * nx + x]) (*bits)[(size_t)y * wpr + (x >> 6)] |= 1ull << (x & 63);
}

double ic_density(const JobSpec& s, int c, int x, int y) {
    IcParams p{s.nx, s.ny, s.ic, s.rho_hi, s.rho_lo, s.radius, s.noise, s.ic_seed};
    return ic_density_at(p, c, x, y);
}

// ---------------------------------------------------------------------------
// Scheduler
// ---------------------------------------------------------------------------
namespace {
// Timed waits use a system-clock deadline (pthread_cond_timedwait, which
// thread sanitizers see); every wait re-checks its condition, so a step of
// the system clock only moves a wake-up.
std::chrono::system_clock::time_point deadline(double ms) {
    return std::chrono::system_clock::now() +
           std::chrono::duration_cast<std::chrono::system_clock::duration>(std::chrono::duration<double, std::milli>(ms));
}
}  // namespace

// Watchdog tolerance by storage form.  FP64 and the FP32 deviation form
// conserve mass to their rounding; FP16 storage rounds each population to 11
// bits of (f_i - w_i rho0) * 2^8, which has moved the mass of a droplet
// lattice (densities 1 and 0.06, rho0 = 1) by 1e-3 in 144 steps.  The
// watchdog is there for blow-ups (non-finite or runaway mass).
double default_mass_tol(int precision) {
    switch (precision) {
        case SCL_FP64: return 1e-6;
        case SCL_FP32_DEV: return 1e-3;
        default: return 0.1;
    }
}

Scheduler::Scheduler(std::unique_ptr<Backend> be, const scl_server_config& cfg)
    : be_(std::move(be)), cfg_(cfg), arena_(cfg.budget_bytes, 4096) {
    if (cfg_.mass_tol <= 0) cfg_.mass_tol = default_mass_tol(cfg_.precision);
    t0_ = std::chrono::steady_clock::now();
    thr_ = std::thread(&Scheduler::loop, this);
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned n = std::max(1u, std::min(4u, hw / 2));
    for (unsigned k = 0; k < n; ++k) prep_.emplace_back(&Scheduler::prep_loop, this);
}

Scheduler::~Scheduler() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    cv_prep_.notify_all();
    for (std::thread& t : prep_) t.join();
    if (thr_.joinable()) thr_.join();
}

// Mask builders: host work that precedes admission, done off the request
// loop so the device keeps running rounds meanwhile: a caller's mask is
// packed here; a generated one is only sampled (the disks' centres and
// sizes, no rasterising), and the backend rasterises it on the device when
// it builds the job.  Jobs are prepared in release order.
void Scheduler::prep_loop() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        cv_prep_.wait(lk, [&] { return stop_ || !prep_q_.empty(); });
        if (stop_) return;
        const long long id = prep_q_.begin()->second;
        prep_q_.erase(prep_q_.begin());
        auto it = jobs_.find(id);
        if (it == jobs_.end() || finished(it->second->state)) continue;
        const JobSpec& src = it->second->spec;
        JobSpec g;   // the geometry part of the spec
        g.nx = src.nx;
        g.ny = src.ny;
        g.mask = src.mask;
        g.porosity = src.porosity;
        g.rmin = src.rmin;
        g.rmax = src.rmax;
        g.walls = src.walls;
        g.geom_seed = src.geom_seed;
        lk.unlock();
        std::vector<uint64_t> bits;
        std::vector<Disk> disks;
        long long nf = 0;
        const bool device_mask = g.mask.empty() && be_->device_masks();
        if (device_mask) {
            nf = job_disks(g, &disks);
        } else {
            std::vector<int> m;
            job_mask(g, &m);
            pack_mask(m, g.nx, g.ny, &bits);
            for (uint64_t w : bits) nf += __builtin_popcountll(w);
        }
        lk.lock();
        it = jobs_.find(id);
        if (it != jobs_.end() && !finished(it->second->state)) {
            it->second->bits.swap(bits);
            it->second->disks.swap(disks);
            it->second->device_mask = device_mask;
            it->second->fluid_sites = nf;
            it->second->prepared = true;
            cv_work_.notify_one();
        }
    }
}

double Scheduler::now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
}

int Scheduler::submit(const scl_job* job, long long* id) {
    auto j = std::unique_ptr<Job>(new Job());
    std::string err;
    int rc = spec_from_job(job, &j->spec, &err);
    if (rc == SCL_OK) rc = be_->validate(j->spec, &err);
    if (rc == SCL_OK) {
        j->bytes = be_->footprint(j->spec);
        if (arena_.rounded(j->bytes) > arena_.size()) {
            rc = SCL_ERR_RANGE;
            err = "job needs " + std::to_string(j->bytes) + " bytes of device memory; the server's arena holds " +
                  std::to_string(arena_.size());
        }
    }
    std::lock_guard<std::mutex> lk(mu_);
    if (rc != SCL_OK) {
        ++n_rejected_;
        return report(rc, err);
    }
    if ((int)pending_.size() >= cfg_.max_pending) {
        ++n_busy_;
        return report(SCL_ERR_BUSY, "server queue full (" + std::to_string(cfg_.max_pending) + " pending jobs)");
    }
    if (stop_) return report(SCL_ERR_ARG, "server is shutting down");
    j->id = next_id_++;
    j->submit_ms = now_ms();
    Job* p = j.get();
    enqueue_locked(p);
    jobs_[p->id] = std::move(j);
    prep_q_.emplace(p->spec.release_ms, p->id);
    ++n_submitted_;
    if (id) *id = p->id;
    cv_prep_.notify_one();
    cv_work_.notify_one();
    return SCL_OK;
}

// Must hold mu_.  Into the queue in (release, id) order.
void Scheduler::enqueue_locked(Job* p) {
    auto pos = std::upper_bound(pending_.begin(), pending_.end(), p, [](const Job* a, const Job* b) {
        return a->spec.release_ms < b->spec.release_ms ||
               (a->spec.release_ms == b->spec.release_ms && a->id < b->id);
    });
    pending_.insert(pos, p);
}

void Scheduler::fill_status(const Job& j, scl_job_status* st) const {
    std::memset(st, 0, sizeof *st);
    st->state = j.state;
    st->steps_done = j.steps_done;
    st->release_ms = j.spec.release_ms;
    st->admit_ms = j.admit_ms;
    st->done_ms = j.done_ms;
    for (int c = 0; c < 2; ++c) {
        st->mass[c] = j.mass[c];
        st->mass0[c] = j.mass0[c];
    }
    st->bytes = j.bytes;
    std::strncpy(st->error, j.error.c_str(), sizeof st->error - 1);
}

int Scheduler::status(long long id, scl_job_status* st) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return report(SCL_ERR_ARG, "unknown job id " + std::to_string(id));
    if (st) fill_status(*it->second, st);
    return SCL_OK;
}

int Scheduler::wait(long long id, double timeout_ms, scl_job_status* st) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return report(SCL_ERR_ARG, "unknown job id " + std::to_string(id));
    Job* j = it->second.get();
    auto done = [&] { return finished(j->state); };
    bool ok;
    if (timeout_ms < 0) {
        cv_done_.wait(lk, done);
        ok = true;
    } else {
        ok = cv_done_.wait_until(lk, deadline(timeout_ms), done);
    }
    if (st) fill_status(*j, st);
    if (!ok) return report(SCL_ERR_TIMEOUT, "job " + std::to_string(id) + " not finished within the wait");
    return SCL_OK;
}

int Scheduler::cancel(long long id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return report(SCL_ERR_ARG, "unknown job id " + std::to_string(id));
    Job* j = it->second.get();
    if (finished(j->state)) return SCL_OK;
    j->cancel = true;
    if (j->state == SCL_JOB_SUSPENDED) {
        j->saved.clear();
        retire(j, SCL_JOB_CANCELLED, "cancelled while suspended");
        return SCL_OK;
    }
    auto pos = std::find(pending_.begin(), pending_.end(), j);
    if (pos != pending_.end()) {
        pending_.erase(pos);
        j->state = SCL_JOB_CANCELLED;
        j->done_ms = now_ms();
        ++n_cancelled_;
        cv_done_.notify_all();
    }
    cv_work_.notify_one();
    return SCL_OK;
}

int Scheduler::suspend(long long id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return report(SCL_ERR_ARG, "unknown job id " + std::to_string(id));
    Job* j = it->second.get();
    if (finished(j->state)) return report(SCL_ERR_ARG, "job " + std::to_string(id) + " has finished");
    if (j->state == SCL_JOB_SUSPENDED) return SCL_OK;
    if (!be_->can_save()) return report(SCL_ERR_UNSUPPORTED, "this server's backend cannot pre-empt jobs");
    auto pos = std::find(pending_.begin(), pending_.end(), j);
    if (pos != pending_.end()) {   // queued: set aside now (a resumed job keeps its saved state)
        pending_.erase(pos);
        j->state = SCL_JOB_SUSPENDED;
        cv_done_.notify_all();
        return SCL_OK;
    }
    j->suspend = true;   // resident or being set up: at the next round boundary
    cv_work_.notify_one();
    return SCL_OK;
}

int Scheduler::resume(long long id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return report(SCL_ERR_ARG, "unknown job id " + std::to_string(id));
    Job* j = it->second.get();
    if (j->state != SCL_JOB_SUSPENDED) {
        if (j->suspend && !finished(j->state)) {   // not yet off the device: keep running
            j->suspend = false;
            return SCL_OK;
        }
        return report(SCL_ERR_ARG, "job " + std::to_string(id) + " is not suspended");
    }
    if (stop_) return report(SCL_ERR_ARG, "server is shutting down");
    j->state = SCL_JOB_QUEUED;
    enqueue_locked(j);
    cv_work_.notify_one();
    return SCL_OK;
}

int Scheduler::result(long long id, int which, int comp, double* out, size_t n) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return report(SCL_ERR_ARG, "unknown job id " + std::to_string(id));
    const Job& j = *it->second;
    if (j.state != SCL_JOB_DONE) return report(SCL_ERR_ARG, "job " + std::to_string(id) + " has not completed");
    if (comp < 0 || comp > 1) return report(SCL_ERR_ARG, "component must be 0 or 1");
    const std::vector<double>* v = nullptr;
    if (which == SCL_RHO) v = &j.rho[comp];
    else if (which == SCL_F) v = &j.f[comp];
    else return report(SCL_ERR_ARG, "results are SCL_RHO or SCL_F");
    if (v->empty()) return report(SCL_ERR_ARG, "that result was not requested in the job's deliver bits");
    if (!out || n != v->size())
        return report(SCL_ERR_ARG, "result holds " + std::to_string(v->size()) + " values, buffer " +
                                       std::to_string(n));
    std::memcpy(out, v->data(), n * sizeof(double));
    return SCL_OK;
}

int Scheduler::release(long long id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return report(SCL_ERR_ARG, "unknown job id " + std::to_string(id));
    if (!finished(it->second->state)) return report(SCL_ERR_ARG, "job " + std::to_string(id) + " is still live");
    jobs_.erase(it);
    return SCL_OK;
}

int Scheduler::drain(double timeout_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto idle = [&] { return pending_.empty() && active_.empty() && setting_up_ == 0; };
    bool ok;
    if (timeout_ms < 0) {
        cv_done_.wait(lk, idle);
        ok = true;
    } else {
        ok = cv_done_.wait_until(lk, deadline(timeout_ms), idle);
    }
    return ok ? SCL_OK : report(SCL_ERR_TIMEOUT, "jobs still queued or resident at the end of the wait");
}

int Scheduler::stats(scl_server_stats* st) {
    std::lock_guard<std::mutex> lk(mu_);
    std::memset(st, 0, sizeof *st);
    st->submitted = n_submitted_;
    st->admitted = n_admitted_;
    st->done = n_done_;
    st->failed = n_failed_;
    st->cancelled = n_cancelled_;
    st->timed_out = n_timed_out_;
    st->rejected_busy = n_busy_;
    st->rejected = n_rejected_;
    st->rounds = rounds_;
    st->job_steps = job_steps_;
    st->lattice_updates = lups_;
    st->fluid_updates = fluid_lups_;
    st->busy_ms = busy_ms_;
    st->device_ms = be_->device_ms();
    st->setup_ms = setup_ms_;
    st->finish_ms = finish_ms_;
    st->clock_ms = now_ms();
    st->resident = (int)active_.size() + setting_up_;
    st->queued = (int)pending_.size();
    st->max_resident = max_resident_;
    st->arena_bytes = arena_.size();
    st->arena_used = arena_.used();
    st->arena_peak = arena_.peak();
    be_->shape(&st->tile_cols, &st->tile_rows, &st->lookahead);
    st->round_steps = cfg_.round_steps;
    st->workgroups = be_->workgroups();
    st->steps_per_launch = be_->steps_per_launch();
    st->pair_updates = be_->pair_updates();
    return SCL_OK;
}

// Must hold mu_.  Ends a job that is no longer queued: frees its arena block
// and wakes waiters.  The backend state must already be dropped.
void Scheduler::retire(Job* j, int state, const std::string& err) {
    if (j->has_block) {
        arena_.release(j->off);
        j->has_block = false;
    }
    j->state = state;
    j->error = err;
    j->done_ms = now_ms();
    switch (state) {
        case SCL_JOB_DONE: ++n_done_; break;
        case SCL_JOB_FAILED: ++n_failed_; break;
        case SCL_JOB_CANCELLED: ++n_cancelled_; break;
        default: ++n_timed_out_; break;
    }
    cv_done_.notify_all();
}

// Must hold mu_.
void Scheduler::admit_locked(double now, std::vector<Job*>* admitted) {
    int slots = cfg_.max_jobs - (int)active_.size() - setting_up_;
    std::vector<Job*> passed;
    for (auto it = pending_.begin(); it != pending_.end() && slots > 0;) {
        Job* j = *it;
        if (j->spec.release_ms > now) break;
        if (!j->prepared) break;   // its mask is being built: keep release order
        size_t off;
        if (arena_.alloc(j->bytes, &off)) {
            j->off = off;
            j->has_block = true;
            admitted->push_back(j);
            it = pending_.erase(it);
            --slots;
            ++setting_up_;
            ++n_admitted_;
            for (Job* w : passed) ++w->bypassed;
        } else {
            if (j->bypassed >= cfg_.max_bypass) break;   // hold the door for it
            passed.push_back(j);
            ++it;
        }
    }
}

void Scheduler::loop() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        if (stop_) break;
        const double now = now_ms();
        // 1. time-outs and cancellations
        for (auto it = pending_.begin(); it != pending_.end();) {
            Job* j = *it;
            const bool late = j->spec.timeout_ms > 0 && now > j->spec.release_ms + j->spec.timeout_ms;
            if (late) {
                it = pending_.erase(it);
                retire(j, SCL_JOB_TIMED_OUT, "timed out in the queue");
            } else {
                ++it;
            }
        }
        std::vector<std::pair<Job*, int>> evict;
        std::vector<Job*> suspends;
        for (Job* j : active_) {
            if (j->cancel) evict.emplace_back(j, SCL_JOB_CANCELLED);
            else if (j->spec.timeout_ms > 0 && now > j->spec.release_ms + j->spec.timeout_ms)
                evict.emplace_back(j, SCL_JOB_TIMED_OUT);
            else if (j->suspend) suspends.push_back(j);
        }
        // 2. admission
        std::vector<Job*> admitted;
        admit_locked(now, &admitted);
        if (admitted.empty() && active_.empty() && evict.empty() && suspends.empty()) {
            cv_done_.notify_all();   // drain() waiters: nothing resident
            // sleep until the next prepared job is released, or until a
            // submission, a finished preparation, a cancel or shutdown
            double next = -1;
            for (Job* j : pending_)
                if (j->prepared && j->spec.release_ms > now) {
                    next = j->spec.release_ms;
                    break;
                }
            for (Job* j : pending_)   // ... or until a queued job's time-out
                if (j->spec.timeout_ms > 0) {
                    const double dl = j->spec.release_ms + j->spec.timeout_ms;
                    if (next < 0 || dl < next) next = dl;
                }
            if (next < 0) cv_work_.wait(lk);
            else cv_work_.wait_until(lk, deadline(next - now));
            continue;
        }
        lk.unlock();
        for (auto& e : evict) be_->drop(*e.first);
        // pre-emption: the exact state to host memory, then off the device
        std::vector<std::pair<Job*, std::string>> suspend_failed;
        for (Job* j : suspends) {
            std::string err;
            std::vector<char> st;
            const int rc = be_->save(*j, &st, &err);
            be_->drop(*j);
            if (rc != SCL_OK) suspend_failed.emplace_back(j, "suspension failed: " + err);
            else j->saved.swap(st);
        }
        // 3. new lattices
        std::vector<std::pair<Job*, std::string>> setup_failed;
        std::vector<Job*> ready;
        const double ts0 = now_ms();
        for (Job* j : admitted) {
            std::string err;
            const int rc = be_->setup(*j, &err);
            if (rc != SCL_OK) {
                be_->drop(*j);
                setup_failed.emplace_back(j, "setup failed: " + err);
            } else {
                ready.push_back(j);
            }
        }
        const double ts1 = now_ms();
        lk.lock();
        setup_ms_ += ts1 - ts0;
        for (auto& e : evict) {
            active_.erase(std::find(active_.begin(), active_.end(), e.first));
            retire(e.first, e.second, e.second == SCL_JOB_CANCELLED ? "cancelled" : "timed out while resident");
        }
        for (Job* j : suspends) {
            active_.erase(std::find(active_.begin(), active_.end(), j));
            if (std::find_if(suspend_failed.begin(), suspend_failed.end(),
                             [j](const std::pair<Job*, std::string>& e) { return e.first == j; }) !=
                suspend_failed.end())
                continue;
            arena_.release(j->off);
            j->has_block = false;
            j->suspend = false;
            if (j->cancel) {   // cancelled while its state was being saved
                j->saved.clear();
                retire(j, SCL_JOB_CANCELLED, "cancelled");
            } else {
                j->state = SCL_JOB_SUSPENDED;
            }
        }
        for (auto& e : suspend_failed) retire(e.first, SCL_JOB_FAILED, e.second);
        if (!suspends.empty()) cv_done_.notify_all();
        for (auto& e : setup_failed) {
            --setting_up_;
            retire(e.first, SCL_JOB_FAILED, e.second);
        }
        const double t_admit = now_ms();
        for (Job* j : ready) {
            --setting_up_;
            j->state = SCL_JOB_RUNNING;
            j->admit_ms = t_admit;
            active_.push_back(j);
        }
        max_resident_ = std::max(max_resident_, (int)active_.size());
        if (active_.empty()) continue;
        // round
        std::vector<Job*> jobs = active_;
        long long most = 0;
        for (Job* j : jobs) most = std::max(most, j->spec.steps - j->steps_done);
        int rsteps = 1;
        while (rsteps < most && rsteps < cfg_.round_steps) rsteps <<= 1;
        rsteps = std::min(rsteps, cfg_.round_steps);
        std::vector<int> steps(jobs.size());
        for (size_t k = 0; k < jobs.size(); ++k)
            steps[k] = (int)std::min<long long>(rsteps, jobs[k]->spec.steps - jobs[k]->steps_done);
        lk.unlock();
        const double t0 = now_ms();
        std::string err;
        int rc = be_->run_round(jobs, steps, rsteps, &err);
        const double t1 = now_ms();
        // 4. watchdog and completions
        std::vector<std::pair<Job*, std::string>> failed;
        std::vector<Job*> complete;
        double t_finish = 0;
        if (rc != SCL_OK) {
            for (Job* j : jobs) failed.emplace_back(j, "device round failed: " + err);
        } else {
            for (size_t k = 0; k < jobs.size(); ++k) {
                Job* j = jobs[k];
                j->steps_done += steps[k];
                std::string why = j->device_error;
                for (int c = 0; c < 2 && why.empty(); ++c) {
                    if (!std::isfinite(j->mass[c])) {
                        why = "non-finite mass of component " + std::to_string(c) + " after " +
                              std::to_string(j->steps_done) + " steps";
                    } else if (j->mass0_set &&
                               std::fabs(j->mass[c] - j->mass0[c]) > cfg_.mass_tol * std::fabs(j->mass0[c])) {
                        char buf[160];
                        snprintf(buf, sizeof buf, "mass of component %d drifted from %.17g to %.17g after %lld steps",
                                 c, j->mass0[c], j->mass[c], j->steps_done);
                        why = buf;
                    }
                }
                if (!why.empty()) {
                    failed.emplace_back(j, why);
                    continue;
                }
                if (!j->mass0_set) {
                    j->mass0[0] = j->mass[0];
                    j->mass0[1] = j->mass[1];
                    j->mass0_set = true;
                }
                if (j->steps_done >= j->spec.steps) {
                    std::string ferr;
                    const double tf0 = now_ms();
                    if (be_->finish(*j, &ferr) != SCL_OK) failed.emplace_back(j, "delivery failed: " + ferr);
                    else complete.push_back(j);
                    t_finish += now_ms() - tf0;
                }
            }
        }
        for (auto& e : failed) be_->drop(*e.first);
        for (Job* j : complete) be_->drop(*j);
        lk.lock();
        ++rounds_;
        busy_ms_ += t1 - t0;
        finish_ms_ += t_finish;
        if (rc == SCL_OK) {
            for (size_t k = 0; k < jobs.size(); ++k) {
                job_steps_ += steps[k];
                lups_ += jobs[k]->sites() * steps[k];
                fluid_lups_ += jobs[k]->fluid_sites * steps[k];
            }
        }
        for (auto& e : failed) {
            active_.erase(std::find(active_.begin(), active_.end(), e.first));
            retire(e.first, SCL_JOB_FAILED, e.second);
        }
        for (Job* j : complete) {
            active_.erase(std::find(active_.begin(), active_.end(), j));
            retire(j, SCL_JOB_DONE, "");
        }
    }
    // shutdown: everything still live ends cancelled
    std::vector<Job*> live = active_;
    lk.unlock();
    for (Job* j : live) be_->drop(*j);
    lk.lock();
    for (Job* j : live) retire(j, SCL_JOB_CANCELLED, "server destroyed");
    active_.clear();
    for (Job* j : pending_) retire(j, SCL_JOB_CANCELLED, "server destroyed");
    pending_.clear();
    for (auto& e : jobs_)
        if (e.second->state == SCL_JOB_SUSPENDED) retire(e.second.get(), SCL_JOB_CANCELLED, "server destroyed");
}

}  // namespace serve
}  // namespace scl

// ===========================================================================
// C-ABI (the backend-specific scl_server_create lives with its backend)
// ===========================================================================
using scl::serve::report;

extern "C" {

int scl_server_default_config(scl_server_config* cfg) {
    if (!cfg) return report(SCL_ERR_ARG, "null config");
    std::memset(cfg, 0, sizeof *cfg);
    cfg->precision = SCL_FP32_DEV;
    cfg->budget_bytes = (size_t)8 << 30;
    cfg->max_jobs = 64;
    cfg->max_pending = 4096;
    cfg->round_steps = 36;   // a multiple of 2 and 3: no single-step tails in full rounds
    cfg->mass_tol = 0.0;   // by precision
    cfg->max_bypass = 16;
    cfg->use_graphs = 1;
    return SCL_OK;
}

int scl_server_submit(scl_server* s, const scl_job* job, long long* id) {
    if (!s) return report(SCL_ERR_ARG, "null server");
    return s->sched->submit(job, id);
}

int scl_server_status(scl_server* s, long long id, scl_job_status* st) {
    if (!s) return report(SCL_ERR_ARG, "null server");
    return s->sched->status(id, st);
}

int scl_server_wait(scl_server* s, long long id, double timeout_ms, scl_job_status* st) {
    if (!s) return report(SCL_ERR_ARG, "null server");
    return s->sched->wait(id, timeout_ms, st);
}

int scl_server_cancel(scl_server* s, long long id) {
    if (!s) return report(SCL_ERR_ARG, "null server");
    return s->sched->cancel(id);
}

int scl_server_suspend(scl_server* s, long long id) {
    if (!s) return report(SCL_ERR_ARG, "null server");
    return s->sched->suspend(id);
}

int scl_server_resume(scl_server* s, long long id) {
    if (!s) return report(SCL_ERR_ARG, "null server");
    return s->sched->resume(id);
}

int scl_server_result(scl_server* s, long long id, int which, int comp, double* out, size_t n) {
    if (!s) return report(SCL_ERR_ARG, "null server");
    return s->sched->result(id, which, comp, out, n);
}

int scl_server_release(scl_server* s, long long id) {
    if (!s) return report(SCL_ERR_ARG, "null server");
    return s->sched->release(id);
}

int scl_server_drain(scl_server* s, double timeout_ms) {
    if (!s) return report(SCL_ERR_ARG, "null server");
    return s->sched->drain(timeout_ms);
}

int scl_server_get_stats(scl_server* s, scl_server_stats* st) {
    if (!s || !st) return report(SCL_ERR_ARG, "null server or stats");
    return s->sched->stats(st);
}

int scl_server_destroy(scl_server* s) {
    delete s;
    return SCL_OK;
}

int scl_job_mask(const scl_job* job, int* mask, size_t n) {
    scl::serve::JobSpec spec;
    std::string err;
    const int rc = scl::serve::spec_from_job(job, &spec, &err);
    if (rc != SCL_OK) return report(rc, err);
    if (!mask || n != (size_t)spec.nx * spec.ny)
        return report(SCL_ERR_ARG, "mask buffer must hold nx*ny ints");
    std::vector<int> m;
    scl::serve::job_mask(spec, &m);
    std::memcpy(mask, m.data(), n * sizeof(int));
    return SCL_OK;
}

}  // extern "C"
