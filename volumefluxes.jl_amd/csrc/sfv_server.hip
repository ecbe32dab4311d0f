// sfv_server.hip — serving: flood scenarios submitted at any time are
// admitted under an HBM byte budget and advanced together with the batch
// engines (continuous batching).  A worker thread owns every HIP call: it
// admits queued requests in arrival order while their device bytes fit the
// budget, replays the batched step graphs over everything admitted (one
// engine for large grids, one for small ones), and retires each request when
// it reaches its end time: the result is copied to host memory and its
// device buffers go back to the server's pool, which admits the next ones.
//
// Each request's result is bitwise that of running it alone
// (simulate_to_time with the same arguments): it keeps its own dt and time,
// and the batched kernels run the same per-tile code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sinfvm_hip.h"
#include "sfv_engine.hpp"

namespace {

enum { REQ_QUEUED = 0, REQ_RUNNING = 1, REQ_DONE = 2, REQ_FAILED = 3 };

struct Request {
    int64_t id = 0;
    sfv_config cfg{};
    std::vector<double> rain_rates, rain_times, factor, B;
    std::vector<unsigned char> U0;     // interior (nx, ny, 3) in the request's precision
    int64_t nbx = 0, nby = 0;
    double t0 = 0, t_end = 0, maximum_timestep = 0;
    int match_endtime = 1;
    int64_t bytes = 0;
    int status = REQ_QUEUED;
    int error = SFV_OK;
    std::string message;
    double submit_s = -1, admit_s = -1, done_s = -1;
    double t = 0;
    int64_t steps = 0;
    std::vector<unsigned char> result;  // interior state at the end
    sfv_handle* h = nullptr;
    int slot = -1;
    int eng = 0;                        // engine: 0 large grids, 1 small grids
};

}  // namespace

struct sfv_server {
    int precision = SFV_F64;
    int device = 0;
    int64_t budget = 0;
    int64_t in_use = 0, peak = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<int64_t, std::unique_ptr<Request>> requests;
    std::thread worker;
    bool stop = false;
    bool worker_failed = false;
    std::string err, worker_err;
    int64_t next_id = 1;
    int64_t submitted = 0, admitted = 0, completed = 0, failed = 0, replays = 0;
    int64_t cell_updates = 0;
    double busy_s = 0, host_admit_s = 0, host_retire_s = 0;
    sfv::DevicePool pool;              // worker-owned: device blocks, streams, pinned mirrors of finished requests
    std::chrono::steady_clock::time_point start = std::chrono::steady_clock::now();

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    }
    void work();
};

namespace {

thread_local std::string g_server_error;

int sfail(sfv_server* s, int code, const std::string& msg) {
    if (s) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = msg;
    } else {
        g_server_error = msg;
    }
    return code;
}

size_t elem_of(const sfv_config& c) { return c.precision == SFV_F32_STORAGE ? 4 : 8; }

void fill_info(const Request& r, sfv_request_info* info) {
    if (!info) return;
    info->id = r.id;
    info->status = r.status;
    info->error = r.error;
    info->t = r.t;
    info->steps = r.steps;
    info->cell_updates = r.steps * (int64_t)r.cfg.nx * r.cfg.ny;
    info->device_bytes = r.bytes;
    info->submit_s = r.submit_s;
    info->admit_s = r.admit_s;
    info->done_s = r.done_s;
}

// Worker: create the request's handle (buffers from the server's pool) and
// upload its inputs.  The engine takes it later, between replays.
int create_request(sfv::DevicePool* pool, Request* r, std::string& why) {
    sfv_config c = r->cfg;
    c.rain_table_rates = r->rain_rates.empty() ? nullptr : r->rain_rates.data();
    c.rain_table_times = r->rain_times.empty() ? nullptr : r->rain_times.data();
    c.infiltration_factor = r->factor.empty() ? nullptr : r->factor.data();
    sfv_handle* h = nullptr;
    int rc = sfv::create_handle(&c, pool, &h, why);
    if (rc) return rc;
    rc = sfv_set_topography(h, r->B.data(), r->nbx, r->nby);
    if (!rc) rc = sfv_set_state(h, r->U0.data(), SFV_LAYOUT_INTERIOR);
    if (rc) {
        why = sfv_last_error(h);
        sfv_destroy(h);
        return rc;
    }
    r->h = h;
    r->B.clear(); r->B.shrink_to_fit();          // inputs are on the device now
    r->U0.clear(); r->U0.shrink_to_fit();
    return SFV_OK;
}

sfv::RunParams run_params(const Request* r) {
    sfv::RunParams rp;
    rp.t0 = r->t0;
    rp.until_endtime = 1;
    rp.endtime = r->t_end;
    rp.match_endtime = r->match_endtime;
    rp.maximum_timestep = r->maximum_timestep;
    rp.max_steps = 0;
    return rp;
}

}  // namespace

// Grids of at most this many cells run in the small-grid engine.
constexpr int64_t SMALL_GRID_CELLS = 1 << 18;

// The worker keeps two batch engines.  Large grids step in one, replayed once
// per loop (8 steps of every member), and while that replay runs the worker
// downloads finished results and creates the newly admitted requests.  Small
// grids, whose steps are short and latency-bound, step in the other, which
// replays concurrently with the large-grid replay in flight: the large
// engine then leaves one block slot per CU free (Share::HOST) and the small
// one runs one block per CU at raised wave priority (Share::GUEST), so a
// small request that needs many steps advances at close to its solo step
// rate while the large grids fill the rest of the device.  Alone, either
// engine fills the device.  Membership changes happen between replays.
void sfv_server::work() {
    std::string why;
    sfv::BatchEngine engs[2];
    int rc = hipSetDevice(device) == hipSuccess ? SFV_OK : SFV_ERR_HIP;
    for (int e = 0; e < 2 && !rc; ++e) {
        rc = engs[e].init(precision, 64, why);
    }
    if (rc) {
        std::lock_guard<std::mutex> lk(mu);
        worker_failed = true;
        worker_err = why.empty() ? "device set-up failed" : why;
        for (Request* r : queue) { r->status = REQ_FAILED; r->error = rc; r->message = worker_err; ++failed; }
        queue.clear();
        cv_done.notify_all();
        return;
    }
    auto finish = [&](Request* r, int status, int error, const std::string& msg) {
        // caller holds mu
        r->status = status;
        r->error = error;
        r->message = msg;
        r->done_s = now();
        in_use -= r->bytes;
        if (status == REQ_DONE) { ++completed; cell_updates += r->steps * (int64_t)r->cfg.nx * r->cfg.ny; }
        else ++failed;
    };
    std::vector<Request*> run[2];      // members of each engine
    std::vector<Request*> to_retire;   // finished, removed from their engine, result not yet downloaded
    std::vector<Request*> created;     // handle made, not yet in an engine
    bool large_in_flight = false;
    double t_large = 0;                // start of the large-grid replay in flight
    auto fail_all = [&](int e, int code) {
        std::lock_guard<std::mutex> lk(mu);
        for (Request* r : run[e]) {
            engs[e].remove(r->slot);
            sfv_destroy(r->h);
            r->h = nullptr;
            finish(r, REQ_FAILED, code, why);
        }
        run[e].clear();
        cv_done.notify_all();
    };
    // after a replay of engine e: finished requests leave it
    auto collect = [&](int e) {
        std::vector<Request*> still;
        for (Request* r : run[e]) {
            if (!engs[e].finished(r->slot)) { still.push_back(r); continue; }
            const sfv::StepCtl& c = engs[e].ctl(r->slot);
            r->t = c.t;
            r->steps = c.steps;
            r->error = c.nonfinite != 0 ? SFV_ERR_NONFINITE_DT : SFV_OK;
            engs[e].remove(r->slot);
            to_retire.push_back(r);
        }
        run[e].swap(still);
    };
    for (;;) {
        // 1. admissions, in arrival order while the next request fits the budget
        std::vector<Request*> admit;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (!large_in_flight && to_retire.empty() && created.empty() && run[1].empty())
                cv_work.wait(lk, [&] { return stop || !queue.empty() || !run[0].empty() || !run[1].empty(); });
            if (stop) break;
            while (!queue.empty()) {
                Request* r = queue.front();
                if (r->bytes > budget) {
                    queue.pop_front();
                    r->status = REQ_FAILED;
                    r->error = SFV_ERR_BUDGET;
                    r->message = "request needs " + std::to_string(r->bytes) + " device bytes, over the budget of " +
                                 std::to_string(budget);
                    r->done_s = now();
                    ++failed;
                    cv_done.notify_all();
                    continue;
                }
                if (in_use + r->bytes > budget) break;
                queue.pop_front();
                in_use += r->bytes;
                peak = std::max(peak, in_use);
                r->status = REQ_RUNNING;
                r->admit_s = now();
                ++admitted;
                admit.push_back(r);
            }
        }
        // 2. host work, overlapping the replays in flight
        if (!to_retire.empty()) {
            const double t0 = now();
            for (Request* r : to_retire) {
                r->result.resize((size_t)r->cfg.nx * r->cfg.ny * 3 * elem_of(r->cfg));
                const int rc4 = sfv_get_state(r->h, r->result.data(), SFV_LAYOUT_INTERIOR);
                const std::string msg4 = rc4 ? sfv_last_error(r->h) : "";
                sfv_destroy(r->h);                   // buffers go back to the pool
                r->h = nullptr;
                std::lock_guard<std::mutex> lk(mu);
                if (rc4) finish(r, REQ_FAILED, rc4, msg4);
                else if (r->error == SFV_ERR_NONFINITE_DT)
                    finish(r, REQ_DONE, SFV_ERR_NONFINITE_DT, "a time step produced a non-finite dt");
                else finish(r, REQ_DONE, SFV_OK, "");
            }
            to_retire.clear();
            std::lock_guard<std::mutex> lk(mu);
            host_retire_s += now() - t0;
            cv_done.notify_all();
        }
        if (!admit.empty()) {
            const double t0 = now();
            for (Request* r : admit) {
                int64_t held;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    held = in_use;
                }
                // cached blocks and the bytes of admitted requests stay within the budget
                if ((int64_t)pool.cached() + held > budget) pool.trim((size_t)std::max<int64_t>(0, budget - held));
                const int rc2 = create_request(&pool, r, why);
                if (rc2) {
                    std::lock_guard<std::mutex> lk(mu);
                    finish(r, REQ_FAILED, rc2, why);
                    cv_done.notify_all();
                } else {
                    r->eng = (int64_t)r->cfg.nx * r->cfg.ny <= SMALL_GRID_CELLS ? 1 : 0;
                    created.push_back(r);
                }
            }
            std::lock_guard<std::mutex> lk(mu);
            host_admit_s += now() - t0;
        }
        // 3. end the large-grid replay: once it has finished while small grids
        //    step beside it, else by waiting for it
        if (large_in_flight && (run[1].empty() || engs[0].replay_ready())) {
            large_in_flight = false;
            const int rc3 = engs[0].replay_end(why);
            const double dt = now() - t_large;
            {
                std::lock_guard<std::mutex> lk(mu);
                ++replays;
                busy_s += dt;
            }
            if (rc3) fail_all(0, rc3);
            else collect(0);
        }
        // 4. the created requests join their engine while it is idle
        {
            std::vector<Request*> later;
            for (Request* r : created) {
                if (r->eng == 0 && large_in_flight) { later.push_back(r); continue; }
                const int rc5 = engs[r->eng].add(r->h, run_params(r), &r->slot, why);
                if (rc5) {
                    sfv_destroy(r->h);
                    r->h = nullptr;
                    std::lock_guard<std::mutex> lk(mu);
                    finish(r, REQ_FAILED, rc5, why);
                    cv_done.notify_all();
                } else {
                    run[r->eng].push_back(r);
                }
            }
            created.swap(later);
        }
        // 5. the next large-grid replay, left in flight
        if (!large_in_flight && !run[0].empty()) {
            engs[0].set_share(run[1].empty() ? sfv::Share::ALONE : sfv::Share::HOST);
            t_large = now();
            const int rc8 = engs[0].replay_begin(1, why);
            if (rc8) {
                (void)engs[0].replay_end(why);
                fail_all(0, rc8);
            } else {
                large_in_flight = true;
            }
        }
        // 6. small grids: 2 replays (16 steps) beside the large-grid replay,
        //    or 4 alone, then their finished requests leave
        if (!run[1].empty()) {
            const bool beside = large_in_flight;
            engs[1].set_share(beside ? sfv::Share::GUEST : sfv::Share::ALONE);
            const int k = beside ? 2 : 4;
            const double t0 = now();
            int rc6 = engs[1].replay_begin(k, why);
            const int rc7 = engs[1].replay_end(why);
            if (!rc6) rc6 = rc7;
            {
                std::lock_guard<std::mutex> lk(mu);
                replays += k;
                if (!beside) busy_s += now() - t0;
            }
            if (rc6) fail_all(1, rc6);
            else collect(1);
        }
    }
    // stopped: release what is still on the device
    if (large_in_flight) (void)engs[0].replay_end(why);
    std::lock_guard<std::mutex> lk(mu);
    for (int e = 0; e < 2; ++e) {
        for (Request* r : run[e]) {
            engs[e].remove(r->slot);
            sfv_destroy(r->h);
            r->h = nullptr;
            finish(r, REQ_FAILED, SFV_ERR_STATE, "server stopped");
        }
        run[e].clear();
    }
    for (Request* r : to_retire) {
        sfv_destroy(r->h);
        r->h = nullptr;
        finish(r, REQ_FAILED, SFV_ERR_STATE, "server stopped");
    }
    to_retire.clear();
    for (Request* r : created) {
        sfv_destroy(r->h);
        r->h = nullptr;
        finish(r, REQ_FAILED, SFV_ERR_STATE, "server stopped");
    }
    created.clear();
    for (Request* r : queue) {
        r->status = REQ_FAILED; r->error = SFV_ERR_STATE; r->message = "server stopped"; ++failed;
    }
    queue.clear();
    cv_done.notify_all();
}

extern "C" {

int sfv_server_create(int64_t budget_bytes, int32_t precision, sfv_server** out) {
    if (!out) return sfail(nullptr, SFV_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (budget_bytes <= 0) return sfail(nullptr, SFV_ERR_ARGUMENT, "budget_bytes must be positive");
    if (precision != SFV_F64 && precision != SFV_F32_STORAGE) return sfail(nullptr, SFV_ERR_ARGUMENT, "unknown precision");
    int dev = 0, n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return sfail(nullptr, SFV_ERR_HIP, "no HIP device");
    if (hipGetDevice(&dev) != hipSuccess) return sfail(nullptr, SFV_ERR_HIP, "hipGetDevice failed");
    sfv_server* s = new sfv_server();
    s->precision = precision;
    s->budget = budget_bytes;
    s->device = dev;
    s->worker = std::thread([s] { s->work(); });
    *out = s;
    return SFV_OK;
}

void sfv_server_destroy(sfv_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    delete s;
}

const char* sfv_server_last_error(const sfv_server* s) {
    return s ? s->err.c_str() : g_server_error.c_str();
}

int sfv_server_submit(sfv_server* s, const sfv_config* cfg, const double* B, int64_t nbx, int64_t nby,
                      const void* U_interior, double t0, double t_end, int32_t match_endtime,
                      double maximum_timestep, int64_t* id_out) {
    if (!s || !cfg || !B || !U_interior || !id_out) return sfail(s, SFV_ERR_ARGUMENT, "NULL argument");
    if (cfg->precision != s->precision) return sfail(s, SFV_ERR_ARGUMENT, "request precision differs from the server's");
    if (cfg->stepper != SFV_RUNGE_KUTTA2) return sfail(s, SFV_ERR_ARGUMENT, "served requests use RungeKutta2");
    if (cfg->rain == SFV_RAIN_FIELD) return sfail(s, SFV_ERR_ARGUMENT, "FunctionalRain cannot be served");
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->gc < 1 || cfg->gc > sfv::MAX_GC) return sfail(s, SFV_ERR_DOMAIN, "bad grid size");
    const int64_t nxT = cfg->nx + 2 * (int64_t)cfg->gc, nyT = cfg->ny + 2 * (int64_t)cfg->gc;
    if (nbx != nxT + 1 || nby != nyT + 1)
        return sfail(s, SFV_ERR_DOMAIN, "Bottom topography should be of size (" + std::to_string(nxT + 1) + ", " +
                                            std::to_string(nyT + 1) + ")");
    {   // the fused kernels need every ghost source inside the grid
        sfv_handle probe;
        probe.cfg = *cfg;
        probe.build_params();
        if (!probe.fusable())
            return sfail(s, SFV_ERR_ARGUMENT, "served requests need ghost sources inside the grid "
                                              "(WallBC or PeriodicBC with nx, ny >= gc)");
    }
    std::unique_ptr<Request> r(new Request());
    r->cfg = *cfg;
    if (cfg->rain == SFV_RAIN_TIME_TABLE && cfg->rain_table_len > 0 && cfg->rain_table_rates && cfg->rain_table_times) {
        r->rain_rates.assign(cfg->rain_table_rates, cfg->rain_table_rates + cfg->rain_table_len);
        r->rain_times.assign(cfg->rain_table_times, cfg->rain_table_times + cfg->rain_table_len);
    }
    if (cfg->infiltration_factor) r->factor.assign(cfg->infiltration_factor, cfg->infiltration_factor + nxT * nyT);
    r->cfg.rain_table_rates = r->cfg.rain_table_times = nullptr;
    r->cfg.infiltration_factor = nullptr;
    r->B.assign(B, B + nbx * nby);
    r->nbx = nbx;
    r->nby = nby;
    const size_t ubytes = (size_t)cfg->nx * cfg->ny * 3 * elem_of(*cfg);
    r->U0.assign((const unsigned char*)U_interior, (const unsigned char*)U_interior + ubytes);
    r->t0 = t0;
    r->t = t0;
    r->t_end = t_end;
    r->match_endtime = match_endtime ? 1 : 0;
    r->maximum_timestep = maximum_timestep;
    r->bytes = sfv_handle::init_bytes(*cfg, true);   // pooled blocks are class-rounded
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop || s->worker_failed) {
        s->err = s->worker_failed ? s->worker_err : "server stopped";
        return SFV_ERR_STATE;
    }
    r->id = s->next_id++;
    r->submit_s = s->now();
    *id_out = r->id;
    s->queue.push_back(r.get());
    s->requests[r->id] = std::move(r);
    ++s->submitted;
    s->cv_work.notify_one();
    return SFV_OK;
}

int sfv_server_poll(sfv_server* s, int64_t id, sfv_request_info* info) {
    if (!s) return SFV_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) {
        s->err = "unknown request id";
        return SFV_ERR_ARGUMENT;
    }
    fill_info(*it->second, info);
    return SFV_OK;
}

int sfv_server_wait(sfv_server* s, int64_t id, double timeout_s, sfv_request_info* info) {
    if (!s) return SFV_ERR_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) {
        s->err = "unknown request id";
        return SFV_ERR_ARGUMENT;
    }
    Request* r = it->second.get();
    auto ready = [&] { return r->status == REQ_DONE || r->status == REQ_FAILED; };
    if (timeout_s < 0) s->cv_done.wait(lk, ready);
    else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
    fill_info(*r, info);
    if (!ready()) {
        s->err = "timed out";
        return SFV_ERR_STATE;
    }
    return SFV_OK;
}

int sfv_server_fetch(sfv_server* s, int64_t id, void* U_interior_out, sfv_request_info* info) {
    if (!s) return SFV_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) {
        s->err = "unknown request id";
        return SFV_ERR_ARGUMENT;
    }
    Request* r = it->second.get();
    fill_info(*r, info);
    if (r->status == REQ_FAILED) {
        s->err = r->message;
        const int code = r->error ? r->error : SFV_ERR_STATE;
        s->requests.erase(it);
        return code;
    }
    if (r->status != REQ_DONE) {
        s->err = "request not finished";
        return SFV_ERR_STATE;
    }
    if (U_interior_out) std::memcpy(U_interior_out, r->result.data(), r->result.size());
    const int code = r->error;   // SFV_OK or SFV_ERR_NONFINITE_DT (result still delivered)
    if (code) s->err = r->message;
    s->requests.erase(it);
    return code;
}

int sfv_server_get_stats(sfv_server* s, sfv_server_stats* out) {
    if (!s || !out) return SFV_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->submitted;
    out->admitted = s->admitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->queued = (int64_t)s->queue.size();
    int64_t running = 0;
    for (auto& kv : s->requests) running += kv.second->status == REQ_RUNNING;
    out->running = running;
    out->bytes_in_use = s->in_use;
    out->bytes_peak = s->peak;
    out->budget_bytes = s->budget;
    out->cell_updates = s->cell_updates;
    out->replays = s->replays;
    out->busy_s = s->busy_s;
    out->uptime_s = s->now();
    out->host_admit_s = s->host_admit_s;
    out->host_retire_s = s->host_retire_s;
    return SFV_OK;
}

}  // extern "C"
