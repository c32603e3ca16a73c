// serve_stress.cpp -- TEST INFRASTRUCTURE.  Concurrent clients against the
// serving scheduler (CPU backend, tests/drivers/serve_cpu.cpp): submissions
// with back-pressure retries, release times, cancellations, pre-emptions
// (suspend and resume), waits, result reads, releases and stats polling from
// several threads at once.  Built with
// ThreadSanitizer on the host compiler; exits non-zero if any job ends in an
// unexpected state or a delivered result is not finite.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <thread>
#include <vector>

#include "../../include/sc_lbm_amd.h"

extern "C" int sct_server_create(const scl_server_config* cfg, scl_server** out);

int main() {
    scl_server_config cfg;
    scl_server_default_config(&cfg);
    cfg.precision = SCL_FP64;
    cfg.budget_bytes = (size_t)3 * (32 * 32 * 192 + 4096);   // about three resident jobs
    cfg.max_jobs = 4;
    cfg.max_pending = 16;
    cfg.round_steps = 4;
    scl_server* srv = nullptr;
    if (sct_server_create(&cfg, &srv) != SCL_OK) return 2;
    std::atomic<int> bad{0}, done{0}, cancelled{0}, busy{0};
    std::atomic<bool> stop{false};
    std::thread poller([&] {
        while (!stop) {
            scl_server_stats st;
            scl_server_get_stats(srv, &st);
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    });
    auto client = [&](int t) {
        for (int k = 0; k < 24; ++k) {
            scl_job j = {};
            j.nx = 16 + 8 * ((t + k) % 3);
            j.ny = 16 + 8 * (k % 3);
            j.porosity = 1.0;
            j.ic = SCL_IC_DROPLET;
            j.rho_hi = 1.0;
            j.rho_lo = 0.1;
            j.radius = 5;
            j.noise = 0.01;
            j.ic_seed = (uint64_t)(t * 100 + k);
            j.tau[0] = j.tau[1] = 1.0;
            j.g_fluid[0] = j.g_fluid[1] = 0.9;
            j.steps = 3 + (k * 7) % 20;
            j.deliver = (k % 2) ? SCL_DELIVER_RHO : 0;
            long long id = 0;
            int rc;
            while ((rc = scl_server_submit(srv, &j, &id)) == SCL_ERR_BUSY) {
                ++busy;
                std::this_thread::sleep_for(std::chrono::microseconds(300));
            }
            if (rc != SCL_OK) {
                ++bad;
                continue;
            }
            if (k % 5 == 3) scl_server_cancel(srv, id);
            if (k % 5 == 1) {
                // pre-empt and resume; a job that already finished refuses both
                scl_server_suspend(srv, id);
                std::this_thread::sleep_for(std::chrono::microseconds(400));
                scl_server_resume(srv, id);
            }
            scl_job_status st;
            if (scl_server_wait(srv, id, 30000.0, &st) != SCL_OK) {
                ++bad;
                continue;
            }
            if (st.state == SCL_JOB_DONE) {
                ++done;
                if (j.deliver) {
                    std::vector<double> rho((size_t)j.nx * j.ny);
                    if (scl_server_result(srv, id, SCL_RHO, 0, rho.data(), rho.size()) != SCL_OK) ++bad;
                    for (double v : rho)
                        if (!std::isfinite(v)) ++bad;
                }
            } else if (st.state == SCL_JOB_CANCELLED && k % 5 == 3) {
                ++cancelled;
            } else {
                ++bad;
            }
            scl_server_release(srv, id);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < 4; ++t) th.emplace_back(client, t);
    for (auto& x : th) x.join();
    stop = true;
    poller.join();
    scl_server_drain(srv, 30000.0);
    scl_server_destroy(srv);
    std::printf("done %d cancelled %d busy-retries %d bad %d\n", done.load(), cancelled.load(), busy.load(), bad.load());
    return bad.load() == 0 && done.load() + cancelled.load() == 96 ? 0 : 1;
}
