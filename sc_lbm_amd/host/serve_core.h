// serve_core.h -- device-independent core of the serving layer: job records,
// the HBM arena allocator, the job generator and the scheduler (request
// loop, admission by memory budget, rounds of steps, watchdog, cancellation
// and time-outs).  A Backend owns the lattices of the resident jobs and
// advances them; libsclbm.so pairs the scheduler with the gfx950 backend.
//
// Plain C++14: no HIP types, so the scheduler can also be exercised against
// a CPU backend.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sc_lbm_amd.h"

namespace scl {
namespace serve {

// Records msg as the calling thread's last error (scl_last_error) and
// returns code.  Defined by the library that links the scheduler.
int report(int code, const std::string& msg);

// ---------------------------------------------------------------------------
// Arena: first-fit allocator over [0, size) with coalescing free list.
// Offsets are multiples of `align`.
// ---------------------------------------------------------------------------
class Arena {
  public:
    Arena(size_t size, size_t align);
    bool alloc(size_t bytes, size_t* off);
    void release(size_t off);
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t largest_free() const;
    size_t rounded(size_t bytes) const { return (bytes + align_ - 1) / align_ * align_; }

  private:
    size_t size_, align_, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    std::map<size_t, size_t> live_;   // offset -> length
};

// ---------------------------------------------------------------------------
// Job specification (an owned copy of scl_job) and the generators behind its
// defaults.
// ---------------------------------------------------------------------------
struct JobSpec {
    int nx = 0, ny = 0;
    std::vector<int> mask;            // empty: generated
    double porosity = 1.0;
    int rmin = 8, rmax = 20, walls = 0;
    uint64_t geom_seed = 0;
    std::vector<double> f[2];         // empty: generated
    int ic = SCL_IC_UNIFORM;
    double rho_hi = 1.0, rho_lo = 1.0, radius = 0.0, noise = 0.0;
    uint64_t ic_seed = 0;
    double tau[2] = {1.0, 1.0}, g_fluid[2] = {0.0, 0.0}, g_solid[2] = {0.0, 0.0};
    bool has_force[2] = {false, false};
    double force[2][9] = {};
    long long steps = 0;
    double release_ms = 0.0, timeout_ms = 0.0;
    int deliver = 0;
    double rho0[2] = {1.0, 1.0};      // deviation reference density (resolved, 2^-10 multiples)
};

// Validates and copies a caller's scl_job.  Returns SCL_OK or an error with
// the message in *err.
int spec_from_job(const scl_job* j, JobSpec* out, std::string* err);

// The mask of a job, 1 = fluid, nx*ny ints: the caller's, or the seeded
// porous sampler (non-overlapping disks of integer radius in [rmin, rmax],
// add_circle node sets, optional solid rows y = 0 and y = ny-1, placed by
// rejection until the fluid fraction is at most `porosity`).
void job_mask(const JobSpec& s, std::vector<int>* mask);

// An add_circle call of the sampler: centre (xc, yc), odd diameter L.
struct Disk {
    int xc, yc, L;
};

// The generated geometry of a job (no caller mask) without rasterising it:
// the sampler's disks in placement order, and the fluid sites of the mask
// they and the walls leave (the disks never overlap each other or the walls,
// so each removes exactly its node count).  job_mask is these disks carved
// from the walled lattice.
long long job_disks(const JobSpec& s, std::vector<Disk>* disks);

// Nodes of an add_circle node set of odd diameter L.
long long disk_nodes(long long L);

// Bit-packed mask: row-major, ceil(nx/64) 64-bit words per row, bit x & 63
// of word x >> 6 is site x (1 = fluid).
void pack_mask(const std::vector<int>& m, int nx, int ny, std::vector<uint64_t>* bits);

// Initial density of component c at site (x, y) for a generated initial
// condition (the device initialiser evaluates the same expression).
double ic_density(const JobSpec& s, int c, int x, int y);

// Default relative mass drift that fails a job, by storage form.
double default_mass_tol(int precision);

// ---------------------------------------------------------------------------
// Job record
// ---------------------------------------------------------------------------
struct Job {
    long long id = 0;
    JobSpec spec;
    int state = SCL_JOB_QUEUED;
    double submit_ms = 0, admit_ms = -1, done_ms = -1;
    long long steps_done = 0;
    int bypassed = 0;                 // later jobs admitted while this one waited
    size_t off = 0, bytes = 0;        // arena block
    bool has_block = false;
    double mass0[2] = {0, 0}, mass[2] = {0, 0};
    bool mass0_set = false;
    bool cancel = false;
    bool suspend = false;             // leave the device at the next round boundary
    std::vector<char> saved;          // exact state of a suspended job (backend format), else empty
    std::string error;
    std::string device_error;         // set by the backend's round: fails the job at the watchdog
    std::vector<double> rho[2], f[2]; // deliverables (reference layout)
    std::vector<uint64_t> bits;       // packed mask (pack_mask) of a caller's mask, built before admission
    std::vector<Disk> disks;          // or: the sampler's disks, rasterised by the backend (device_mask)
    bool device_mask = false;
    bool prepared = false;
    long long fluid_sites = 0;        // fluid sites of the mask (set with bits or disks)
    void* dev = nullptr;              // backend state
    long long sites() const { return (long long)spec.nx * spec.ny; }
};

// ---------------------------------------------------------------------------
// Backend: owns device memory of the resident jobs.
// ---------------------------------------------------------------------------
class Backend {
  public:
    virtual ~Backend() {}
    // Whether the backend can run the job at all (SCL_OK, or an error code
    // with the message in *err).
    virtual int validate(const JobSpec& s, std::string* err) const {
        (void)s;
        (void)err;
        return SCL_OK;
    }
    // Memory a job needs inside the arena.
    virtual size_t footprint(const JobSpec& s) const = 0;
    // Build the job's lattice at arena offset j.off: from j.saved when it is
    // not empty (a resumed job), else from the spec.  On failure return an
    // error code with the message in *err (the job fails, others continue).
    virtual int setup(Job& j, std::string* err) = 0;
    // Copy a resident job's exact state to *out (restored by setup), for
    // pre-emption; backends that cannot, leave the default.
    virtual bool can_save() const { return false; }
    virtual int save(Job& j, std::vector<char>* out, std::string* err) {
        (void)j;
        (void)out;
        *err = "this backend cannot pre-empt jobs";
        return SCL_ERR_UNSUPPORTED;
    }
    // Advance jobs[k] by steps[k] >= 1 steps (all steps[k] <= rsteps); on
    // return jobs[k]->mass holds the FP64 mass per component afterwards.
    virtual int run_round(const std::vector<Job*>& jobs, const std::vector<int>& steps, int rsteps,
                          std::string* err) = 0;
    // Deliverables of a finished job (spec.deliver).
    virtual int finish(Job& j, std::string* err) = 0;
    // Forget the job's device state (its arena block is released by the caller).
    virtual void drop(Job& j) = 0;
    // Launch shape in use (cols, rows, lookahead) and workgroups per step.
    virtual void shape(int* cols, int* rows, int* la) const = 0;
    virtual int workgroups() const { return 0; }
    // Time steps a wide job advances per launch (1 or 2).
    virtual int steps_per_launch() const { return 1; }
    // Device time of the rounds so far, in milliseconds (0 if not measured).
    virtual double device_ms() const { return 0.0; }
    // Lattice updates made by two-step launches so far.
    virtual long long pair_updates() const { return 0; }
    // Whether setup() rasterises a generated mask from Job::disks itself.
    virtual bool device_masks() const { return false; }
};

// ---------------------------------------------------------------------------
// Scheduler
// ---------------------------------------------------------------------------
class Scheduler {
  public:
    Scheduler(std::unique_ptr<Backend> be, const scl_server_config& cfg);
    ~Scheduler();
    int submit(const scl_job* job, long long* id);
    int status(long long id, scl_job_status* st);
    int wait(long long id, double timeout_ms, scl_job_status* st);
    int cancel(long long id);
    int suspend(long long id);
    int resume(long long id);
    int result(long long id, int which, int comp, double* out, size_t n);
    int release(long long id);
    int drain(double timeout_ms);
    int stats(scl_server_stats* st);
    double now_ms() const;
    const scl_server_config& config() const { return cfg_; }

  private:
    void loop();
    void prep_loop();
    void admit_locked(double now, std::vector<Job*>* admitted);
    void retire(Job* j, int state, const std::string& err);
    void fill_status(const Job& j, scl_job_status* st) const;
    bool finished(int state) const { return state >= SCL_JOB_DONE && state <= SCL_JOB_TIMED_OUT; }
    void enqueue_locked(Job* j);

    std::unique_ptr<Backend> be_;
    scl_server_config cfg_;
    Arena arena_;
    std::mutex mu_;
    std::condition_variable cv_work_;   // loop waits for submissions
    std::condition_variable cv_done_;   // callers wait for completions
    std::map<long long, std::unique_ptr<Job>> jobs_;
    std::deque<Job*> pending_;          // in (release, id) order
    std::set<std::pair<double, long long>> prep_q_;   // (release, id) of jobs to prepare
    std::condition_variable cv_prep_;
    std::vector<std::thread> prep_;     // mask builders
    std::vector<Job*> active_;
    long long next_id_ = 1;
    int setting_up_ = 0;                // admitted, lattice being built
    bool stop_ = false;
    std::thread thr_;
    std::chrono::steady_clock::time_point t0_;
    // statistics
    long long n_submitted_ = 0, n_admitted_ = 0, n_done_ = 0, n_failed_ = 0, n_cancelled_ = 0,
              n_timed_out_ = 0, n_busy_ = 0, n_rejected_ = 0, rounds_ = 0, job_steps_ = 0, lups_ = 0,
              fluid_lups_ = 0;
    double busy_ms_ = 0, setup_ms_ = 0, finish_ms_ = 0;
    int max_resident_ = 0;
};

}  // namespace serve
}  // namespace scl

// The C-ABI handle.
struct scl_server {
    std::unique_ptr<scl::serve::Scheduler> sched;
};
