// serve_cpu.cpp -- TEST INFRASTRUCTURE.  The serving layer's scheduler
// (sc_lbm_amd/host/serve_core.cpp, the same source libsclbm.so links) driven
// by a CPU backend that advances each job with the oracle.  Lets the tests
// exercise admission by memory budget, back-pressure, cancellation,
// time-outs, failure isolation and the mass watchdog without a GPU.
//
// Built into tests/drivers/bin/libserve_cpu.so; exports the scl_server_*
// entry points plus sct_server_create and two test hooks.
#include <atomic>
#include <chrono>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../oracle/sc_oracle.h"
#include "../../sc_lbm_amd/host/serve_core.h"

using scl::serve::Job;
using scl::serve::JobSpec;

static thread_local std::string g_err;

namespace scl {
namespace serve {
int report(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
}  // namespace serve
}  // namespace scl

namespace {

std::atomic<int> g_round_delay_ms{0};
std::atomic<unsigned long long> g_fail_seed{~0ull};

class CpuBackend : public scl::serve::Backend {
  public:
    size_t footprint(const JobSpec& s) const override { return (size_t)s.nx * s.ny * 192 + 4096; }

    int setup(Job& j, std::string* err) override {
        const JobSpec& s = j.spec;
        if (s.f[0].empty() && s.ic_seed == g_fail_seed.load()) {
            *err = "injected setup failure";
            return SCL_ERR_HIP;
        }
        // the scheduler's packed mask (the GPU backend uploads the same words)
        const int wpr = (s.nx + 63) / 64;
        if (!j.prepared || j.bits.size() != (size_t)s.ny * wpr) {
            *err = "job admitted without its packed mask";
            return SCL_ERR_ARG;
        }
        std::vector<int> m((size_t)s.nx * s.ny);
        for (int y = 0; y < s.ny; ++y)
            for (int x = 0; x < s.nx; ++x)
                m[(size_t)y * s.nx + x] = (int)((j.bits[(size_t)y * wpr + (x >> 6)] >> (x & 63)) & 1u);
        oracle_sys* o = oracle_create(s.nx, s.ny, 2, m.data(), s.tau, s.g_fluid, s.g_solid);
        if (!o) {
            *err = "oracle_create failed";
            return SCL_ERR_ARG;
        }
        const size_t n = (size_t)s.nx * s.ny;
        for (int c = 0; c < 2; ++c) {
            std::vector<double> f;
            if (!s.f[0].empty()) {
                f = s.f[c];
            } else {
                f.assign(9 * n, 0.0);
                for (int y = 0; y < s.ny; ++y)
                    for (int x = 0; x < s.nx; ++x) {
                        const size_t k = (size_t)y * s.nx + x;
                        if (!m[k]) continue;
                        const double fi = scl::serve::ic_density(s, c, x, y) / 9;
                        for (int i = 0; i < 9; ++i) f[i * n + k] = fi;
                    }
            }
            oracle_set(o, c, ORACLE_F, f.data(), f.size());
        }
        oracle_solid_surface_force(o);
        if (!j.saved.empty()) {   // a resumed job: populations and equilibrium velocities (save)
            if (j.saved.size() != saved_bytes(n)) {
                oracle_destroy(o);
                *err = "saved state does not match the lattice";
                return SCL_ERR_ARG;
            }
            const double* in = (const double*)j.saved.data();
            for (int c = 0; c < 2; ++c, in += 9 * n) oracle_set(o, c, ORACLE_F, in, 9 * n);
            for (int c = 0; c < 2; ++c)
                for (int fld : {ORACLE_UEQX, ORACLE_UEQY}) {
                    oracle_set(o, c, fld, in, n);
                    in += n;
                }
            j.saved.clear();
        }
        j.dev = o;
        return SCL_OK;
    }

    static size_t saved_bytes(size_t n) { return (2 * 9 * n + 4 * n) * sizeof(double); }
    bool can_save() const override { return true; }
    int save(Job& j, std::vector<char>* out, std::string* err) override {
        (void)err;
        const oracle_sys* o = (const oracle_sys*)j.dev;
        const size_t n = (size_t)j.spec.nx * j.spec.ny;
        out->resize(saved_bytes(n));
        double* p = (double*)out->data();
        for (int c = 0; c < 2; ++c, p += 9 * n) oracle_get(o, c, ORACLE_F, p, 9 * n);
        for (int c = 0; c < 2; ++c)
            for (int fld : {ORACLE_UEQX, ORACLE_UEQY}) {
                oracle_get(o, c, fld, p, n);
                p += n;
            }
        return SCL_OK;
    }

    int run_round(const std::vector<Job*>& jobs, const std::vector<int>& steps, int rsteps,
                  std::string* err) override {
        (void)rsteps;
        (void)err;
        const int d = g_round_delay_ms.load();
        if (d > 0) std::this_thread::sleep_for(std::chrono::milliseconds(d));
        for (size_t k = 0; k < jobs.size(); ++k) {
            Job& j = *jobs[k];
            oracle_sys* o = (oracle_sys*)j.dev;
            oracle_run(o, steps[k], j.spec.has_force[0] ? j.spec.force[0] : nullptr,
                       j.spec.has_force[1] ? j.spec.force[1] : nullptr);
            for (int c = 0; c < 2; ++c) {
                double mom[3];
                oracle_moments(o, c, mom);
                j.mass[c] = mom[0];
            }
        }
        return SCL_OK;
    }

    int finish(Job& j, std::string* err) override {
        (void)err;
        oracle_sys* o = (oracle_sys*)j.dev;
        const size_t n = (size_t)j.spec.nx * j.spec.ny;
        for (int c = 0; c < 2; ++c) {
            if (j.spec.deliver & SCL_DELIVER_RHO) {
                oracle_density(o, c);
                j.rho[c].resize(n);
                oracle_get(o, c, ORACLE_RHO, j.rho[c].data(), n);
            }
            if (j.spec.deliver & SCL_DELIVER_F) {
                j.f[c].resize(9 * n);
                oracle_get(o, c, ORACLE_F, j.f[c].data(), 9 * n);
            }
        }
        return SCL_OK;
    }

    void drop(Job& j) override {
        oracle_destroy((oracle_sys*)j.dev);
        j.dev = nullptr;
    }

    void shape(int* cols, int* rows, int* la) const override {
        *cols = 0;
        *rows = 0;
        *la = 0;
    }
};

}  // namespace

extern "C" {

const char* scl_last_error(void) { return g_err.c_str(); }

int sct_server_create(const scl_server_config* cfg, scl_server** out) {
    if (!cfg || !out) return scl::serve::report(SCL_ERR_ARG, "null argument");
    auto s = new scl_server;
    s->sched.reset(new scl::serve::Scheduler(std::unique_ptr<scl::serve::Backend>(new CpuBackend), *cfg));
    *out = s;
    return SCL_OK;
}

// Every round sleeps this long first (to hold jobs resident in a test).
void sct_set_round_delay(int ms) { g_round_delay_ms = ms; }
// Generated jobs with this ic_seed fail their set-up.
void sct_fail_setup_for_seed(unsigned long long seed) { g_fail_seed = seed; }

}  // extern "C"
