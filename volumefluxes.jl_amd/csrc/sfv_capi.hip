// sfv_capi.hip — implementation of the C ABI in include/sinfvm_hip.h:
// handle lifecycle, validation (mirroring the reference's exceptions), state
// transfer, and the orchestration of one time step's kernels on the handle's
// HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/sinfvm_hip.h"
#include "sfv_handle.hpp"

using namespace sfv;

namespace {

thread_local std::string g_create_error;

int fail(sfv_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIP_TRY(h, expr)                                                                             \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return fail((h), SFV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

template <class S>
int step_kernels(sfv_handle* h, int slot, int cur) {
    const KParams& P = h->P;
    hipStream_t st = h->stream;
    S* U0 = (S*)h->U[cur];
    S* U2 = (S*)h->U[cur ^ 1];
    S* RHS = (S*)h->RHS;
    if (h->fused == KV_HYBRID) {
        // the second stage-1 launch also raises this slot's wave-speed maxima,
        // after dt was taken from them; the slot is re-armed before its next use
        if (h->cfg.stepper == SFV_FORWARD_EULER && h->ghost_sources_interior()) {
            // two launches: stage 1 reads ghosts through the boundary maps and
            // the update closes the step; ghosts are filled when the loop ends
            HIP_TRY(h, launch_march<S>(st, P, h->ctl, slot, 1, U0, nullptr, RHS, h->B, 0, 1));
            HIP_TRY(h, launch_finish<S>(st, P, h->ctl, slot, 1, U0, U0, RHS, U2, h->B, 1, 1));
            return SFV_OK;
        }
        if (h->cfg.stepper == SFV_FORWARD_EULER) {
            HIP_TRY(h, launch_march<S>(st, P, h->ctl, slot, 1, U0, nullptr, RHS, h->B));
            HIP_TRY(h, launch_finish<S>(st, P, h->ctl, slot, 1, U0, U0, RHS, U2, h->B, 1));
            HIP_TRY(h, launch_bc<S>(st, P, h->ctl, slot, U2, BC_LAST_STAGE));
            return SFV_OK;
        }
        S* U1 = (S*)h->U[2];
        HIP_TRY(h, launch_march<S>(st, P, h->ctl, slot, 1, U0, nullptr, RHS, h->B));
        HIP_TRY(h, launch_finish<S>(st, P, h->ctl, slot, 1, U0, U0, RHS, U1, h->B, 1));
        HIP_TRY(h, launch_bc<S>(st, P, h->ctl, slot, U1, BC_IN_STEP));
        HIP_TRY(h, launch_march<S>(st, P, h->ctl, slot, 1, U1, nullptr, RHS, h->B));
        HIP_TRY(h, launch_finish<S>(st, P, h->ctl, slot, 2, U0, U1, RHS, U2, h->B, 1));
        HIP_TRY(h, launch_bc<S>(st, P, h->ctl, slot, U2, BC_LAST_STAGE));
        return SFV_OK;
    }
    if (h->cfg.stepper == SFV_FORWARD_EULER) {
        HIP_TRY(h, launch_flux<S>(st, P, h->ctl, slot, U0, RHS, h->B, 1));
        HIP_TRY(h, launch_finish<S>(st, P, h->ctl, slot, 1, U0, U0, RHS, U2, h->B));
        HIP_TRY(h, launch_bc<S>(st, P, h->ctl, slot, U2, BC_LAST_STAGE));
        return SFV_OK;
    }
    if (h->fused == KV_MARCH) {
        // both stages read ghosts through the boundary maps and stage 2 closes
        // the step; the state's ghost cells are filled when the loop ends
        HIP_TRY(h, launch_march<S>(st, P, h->ctl, slot, 1, U0, nullptr, RHS, h->B, 0, 1));
        HIP_TRY(h, launch_march<S>(st, P, h->ctl, slot, 2, U0, RHS, U2, h->B, 1));
        return SFV_OK;
    }
    if (h->fused == KV_TILED) {      // as the march path: ghosts through the maps, stage 2 closes
        HIP_TRY(h, launch_stage1<S>(st, P, h->ctl, slot, U0, RHS, h->B));
        HIP_TRY(h, launch_stage2<S>(st, P, h->ctl, slot, U0, RHS, U2, h->B, 1));
        return SFV_OK;
    }
    S* U1 = (S*)h->U[2];
    HIP_TRY(h, launch_flux<S>(st, P, h->ctl, slot, U0, RHS, h->B, 1));
    HIP_TRY(h, launch_finish<S>(st, P, h->ctl, slot, 1, U0, U0, RHS, U1, h->B));
    HIP_TRY(h, launch_bc<S>(st, P, h->ctl, slot, U1, BC_IN_STEP));
    HIP_TRY(h, launch_flux<S>(st, P, h->ctl, slot, U1, RHS, h->B, 0));
    HIP_TRY(h, launch_finish<S>(st, P, h->ctl, slot, 2, U0, U1, RHS, U2, h->B));
    HIP_TRY(h, launch_bc<S>(st, P, h->ctl, slot, U2, BC_LAST_STAGE));
    return SFV_OK;
}

int launch_steps(sfv_handle* h, int64_t nsteps) {
    if (h->cfg.stepper == SFV_RUNGE_KUTTA2 && (h->fused == KV_PER_PASS || h->fused == KV_HYBRID)) {
        std::string why;
        const int rc = h->ensure_scratch(why);
        if (rc) return fail(h, rc, why);
    }
    for (int64_t s = 0; s < nsteps; ++s) {
        const int slot = (int)((h->step_parity + s) & 1);
        const int cur = (int)((h->cur + s) & 1);
        const int rc = h->cfg.precision == SFV_F64 ? step_kernels<double>(h, slot, cur)
                                                   : step_kernels<float>(h, slot, cur);
        if (rc) return rc;
    }
    return SFV_OK;
}

// Read back the step-control block and account for the steps that ran.
int collect_ctl(sfv_handle* h) {
    HIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return SFV_OK;
}

int ensure_ready(sfv_handle* h) {
    if (!h) return SFV_ERR_ARGUMENT;
    if (!h->has_state) return fail(h, SFV_ERR_STATE, "set the state (set_current_state!) before stepping");
    return SFV_OK;
}

// ---- autotuning of the march kernels' segment height
//
// A wave marches `rows` output rows plus 4 halo rows.  Tall segments waste
// fewer instructions on halo rows, short ones give more waves and a shorter
// serial path per wave, which is what a small grid needs: the best height
// depends on the grid and the device, so it is measured.  Candidates are
// timed on the handle itself (full steps from its current state into the
// spare ping-pong buffer, which the next step overwrites anyway), and the
// winner is cached per device, precision and grid size class (quarter
// octaves of nx and ny).  A march handle whose kernel variant was not set
// with sfv_set_fused also times the LDS-tiled kernels, which step grids of
// roughly 200^2 to 512^2 faster (more waves, each with a shorter serial path),
// and switches to them when they win.  Handles on the hybrid path (ForwardEuler,
// NeumannBC) run only the stage-1 kernel, so they time it alone and keep
// their own cache entries.  Every height gives bitwise the same results.
// SFV_AUTOTUNE=0 in the environment keeps the height model's choice
// (march_rows_per_wave: the fewest rounds of resident workgroups).
struct TuneKey {
    int dev, prec, bx, by, kind;   // kind: 0 full march steps, 1 stage 1 only
    bool operator<(const TuneKey& o) const {
        if (dev != o.dev) return dev < o.dev;
        if (prec != o.prec) return prec < o.prec;
        if (bx != o.bx) return bx < o.bx;
        if (by != o.by) return by < o.by;
        return kind < o.kind;
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tuned;   // the march height; negative: the tiled kernels won (height = its magnitude)

int size_bucket(int n) { return (int)std::lround(4.0 * std::log2((double)std::max(n, 1))); }

template <class S>
hipError_t tune_steps(sfv_handle* h, int slot, int n, bool tiled) {
    const KParams& P = h->P;
    const int cur = (int)(h->cur & 1);
    const S* U0 = (const S*)h->U[cur];
    S* R = (S*)h->RHS;
    S* Us = (S*)h->U[cur ^ 1];
    hipError_t e = hipSuccess;
    for (int k = 0; k < n && e == hipSuccess; ++k) {
        if (tiled) {
            e = launch_stage1<S>(h->stream, P, h->ctl, slot, U0, R, h->B);
            if (e == hipSuccess) e = launch_stage2<S>(h->stream, P, h->ctl, slot, U0, R, Us, h->B);   // not closing
            continue;
        }
        e = launch_march<S>(h->stream, P, h->ctl, slot, 1, U0, nullptr, R, h->B, 0, h->ghost_sources_interior() ? 1 : 0);
        if (h->fused == KV_HYBRID) continue;
        if (e == hipSuccess) e = launch_march<S>(h->stream, P, h->ctl, slot, 2, U0, R, Us, h->B);   // not closing
    }
    return e;
}

}  // namespace

namespace sfv {
void set_call_error(const std::string& msg) { g_create_error = msg; }

// Replay bc.jl's per-ghost assignments (for g = 1..gc, low side then high
// side) on an index map, so each ghost ends up naming the cell whose pre-BC
// value it receives and how many times its normal momentum was negated.
void ghost_maps(int bc, int gc, int N, int* lo, int* hi, int* flo, int* fhi) {
    // Every index the assignments touch lies within gc of an edge's ghost
    // layer: in [0, 2gc) or [N-2gc, N).  The map is kept for those only, in a
    // fixed array (no heap): slot x for x < 2gc, 2gc + (x - (N-2gc)) above.
    constexpr int W = 4 * MAX_GC;
    int cur[W], neg[W];
    const bool direct = N <= W;
    auto at = [&](int x) { return direct || x < 2 * gc ? x : 2 * gc + (x - (N - 2 * gc)); };
    const int span = direct ? N : 4 * gc;
    for (int k = 0; k < span; ++k) neg[k] = 0;
    if (direct) {
        for (int x = 0; x < N; ++x) cur[x] = x;
    } else {
        for (int x = 0; x < 2 * gc; ++x) { cur[at(x)] = x; cur[at(N - 2 * gc + x)] = N - 2 * gc + x; }
    }
    const int f = bc == SFV_BC_WALL ? 1 : 0;
    for (int g = 1; g <= gc; ++g) {
        int dl, sl, dr, sr;
        if (bc == SFV_BC_WALL) {            // bc.jl:92-109
            dl = g - 1; sl = 2 * gc - g; dr = N - gc + g - 1; sr = N - gc - g;
        } else if (bc == SFV_BC_PERIODIC) { // bc.jl:31-39
            dl = g - 1; sl = N + g - 2 * gc - 1; dr = N - (gc - g) - 1; sr = gc + g - 1;
        } else {                            // bc.jl:51-59
            dl = g - 1; sl = 2 * gc - g; dr = N - (gc - g) - 1; sr = N - g;
        }
        cur[at(dl)] = cur[at(sl)]; neg[at(dl)] = neg[at(sl)] ^ f;
        cur[at(dr)] = cur[at(sr)]; neg[at(dr)] = neg[at(sr)] ^ f;
    }
    for (int g = 0; g < gc; ++g) {
        lo[g] = cur[at(g)]; flo[g] = neg[at(g)];
        hi[g] = cur[at(N - gc + g)]; fhi[g] = neg[at(N - gc + g)];
    }
}

int validate_config(const sfv_config* c, std::string& why) {
    if (!c) { why = "config is NULL"; return SFV_ERR_ARGUMENT; }
    if (c->nx < 1 || c->ny < 1) { why = "nx and ny must be >= 1"; return SFV_ERR_DOMAIN; }
    if (c->gc < 1 || c->gc > MAX_GC) { why = "gc must be in [1, 8]"; return SFV_ERR_DOMAIN; }
    // the kernels address within a row by 31-bit byte offsets
    if (c->nx > (1 << 27) - 2 * MAX_GC - 1) { why = "nx must be at most 2^27 - 17"; return SFV_ERR_DOMAIN; }
    if (c->reconstruction == SFV_LINEAR_RECONSTRUCTION && !(c->gc > 1)) {
        why = "LinearReconstruction needs grid.ghostcells > 1 (linearreconstruction.jl:59)";
        return SFV_ERR_ASSERTION;
    }
    if (c->reconstruction != SFV_LINEAR_RECONSTRUCTION && c->reconstruction != SFV_NO_RECONSTRUCTION) {
        why = "unknown reconstruction"; return SFV_ERR_ARGUMENT;
    }
    if (c->bc < SFV_BC_WALL || c->bc > SFV_BC_NEUMANN) { why = "unknown boundary condition"; return SFV_ERR_ARGUMENT; }
    if (c->stepper != SFV_FORWARD_EULER && c->stepper != SFV_RUNGE_KUTTA2) { why = "unknown stepper"; return SFV_ERR_ARGUMENT; }
    if (c->friction < SFV_FRICTION_NONE || c->friction > SFV_FRICTION_FCG2016) { why = "unknown friction law"; return SFV_ERR_ARGUMENT; }
    if (c->rain < SFV_RAIN_NONE || c->rain > SFV_RAIN_FIELD) { why = "unknown rain kind"; return SFV_ERR_ARGUMENT; }
    if (c->rain == SFV_RAIN_TIME_TABLE && (c->rain_table_len < 1 || !c->rain_table_rates || !c->rain_table_times)) {
        why = "TimeDependentRain needs rates and times of equal, non-zero length (source_term_rain.jl:36-38)";
        return SFV_ERR_DOMAIN;
    }
    if (c->infiltration < SFV_INFILTRATION_NONE || c->infiltration > SFV_INFILTRATION_HORTON) {
        why = "unknown infiltration kind"; return SFV_ERR_ARGUMENT;
    }
    if (c->precision != SFV_F64 && c->precision != SFV_F32_STORAGE) { why = "unknown precision"; return SFV_ERR_ARGUMENT; }
    return SFV_OK;
}

}  // namespace sfv

namespace sfv {
bool autotune_enabled() {
    const char* e = std::getenv("SFV_AUTOTUNE");
    return !(e && e[0] == '0');
}

int cached_march_rows(sfv_handle* h) {
    if (h->march_rows > 0) return h->march_rows;
    const TuneKey key{h->device, h->cfg.precision, size_bucket(h->cfg.nx), size_bucket(h->cfg.ny),
                      h->fused == KV_HYBRID ? 1 : 0};
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it == g_tuned.end()) return 0;
    const int rows = it->second < 0 ? -it->second : it->second;
    h->march_rows = rows;
    h->P.rows = rows;
    if (it->second < 0 && !h->variant_forced && h->fused == KV_MARCH) h->fused = KV_TILED;
    return rows;
}

int tune_march_rows(sfv_handle* h, std::string& why) {
    if ((h->fused != KV_MARCH && h->fused != KV_HYBRID) || h->march_rows > 0) return SFV_OK;
    // a march handle whose variant the caller left open also times the tiled kernels
    const bool try_tiled = h->fused == KV_MARCH && !h->variant_forced;
    const int fallback = march_rows_per_wave(h->P, 1, h->cfg.precision == SFV_F64 ? 0 : 1);
    if (!autotune_enabled()) {
        h->march_rows = fallback;
        h->P.rows = fallback;
        return SFV_OK;
    }
    if (cached_march_rows(h) > 0) return SFV_OK;
    const TuneKey key{h->device, h->cfg.precision, size_bucket(h->cfg.nx), size_bucket(h->cfg.ny),
                      h->fused == KV_HYBRID ? 1 : 0};
    std::vector<int> cand;
    for (int c : {1, 2, 3, 4, 6, 8, 12, 16, 24})
        if (c <= h->cfg.ny) cand.push_back(c);
    if (cand.empty() || cand.back() < std::min(h->cfg.ny, 24)) cand.push_back(std::min(h->cfg.ny, 24));
    if (std::find(cand.begin(), cand.end(), fallback) == cand.end()) {   // the model's height is timed too
        cand.push_back(fallback);
        std::sort(cand.begin(), cand.end());
    }
    const int slot = (int)(h->step_parity & 1);
    const bool f64 = h->cfg.precision == SFV_F64;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto bail = [&](hipError_t e) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        h->P.rows = h->march_rows;
        why = std::string("autotuning the march kernels: ") + hipGetErrorString(e);
        return SFV_ERR_HIP;
    };
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess)
        e = launch_prepare(h->stream, h->P, h->ctl, slot, 0.0, NAN, 0, 0, 0.0, -1, 0, h->cfg.precision);
    // time n steps with `rows`; the first call also sizes n so one
    // measurement covers about a millisecond
    auto timed = [&](int rows, int n, double& ms_out, bool tiled = false) -> hipError_t {
        h->P.rows = rows;
        hipError_t r = f64 ? tune_steps<double>(h, slot, 1, tiled) : tune_steps<float>(h, slot, 1, tiled);  // warm-up
        if (r == hipSuccess) r = hipEventRecord(e0, h->stream);
        if (r == hipSuccess) r = f64 ? tune_steps<double>(h, slot, n, tiled) : tune_steps<float>(h, slot, n, tiled);
        if (r == hipSuccess) r = hipEventRecord(e1, h->stream);
        if (r == hipSuccess) r = hipEventSynchronize(e1);
        float ms = 0;
        if (r == hipSuccess) r = hipEventElapsedTime(&ms, e0, e1);
        ms_out = ms / n;
        return r;
    };
    double probe = 0;
    if (e == hipSuccess) e = timed(fallback, 2, probe);
    const int n = (int)std::min(32.0, std::max(3.0, std::ceil(1.0 / std::max(probe, 1e-3))));
    std::vector<double> best(cand.size(), 1e30);
    double best_tiled = 1e30;
    for (int sweep = 0; sweep < 2 && e == hipSuccess; ++sweep) {
        for (size_t c = 0; c < cand.size() && e == hipSuccess; ++c) {
            double ms = 0;
            e = timed(cand[c], n, ms);
            if (e == hipSuccess) best[c] = std::min(best[c], ms);
        }
        if (try_tiled && e == hipSuccess) {
            double ms = 0;
            e = timed(fallback, n, ms, true);
            if (e == hipSuccess) best_tiled = std::min(best_tiled, ms);
        }
    }
    if (e != hipSuccess) return bail(e);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the fastest, unless the model's height is close to it: the timing starts
    // from the current state (often a state at rest), where heights close in
    // time can order differently than in the developed flow.  The model's
    // height is the measured best for grids whose waves fill at least three
    // quarters of the device's resident wave slots (3 per SIMD), so there it
    // must lose by 10% to be replaced; for smaller grids, where the serial
    // path of a wave decides, by 3%.
    int cus = 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess || cus <= 0)
        cus = 256;
    const long waves = (long)((h->cfg.nx + MARCH_COLS - 1) / MARCH_COLS) * ((h->cfg.ny + fallback - 1) / fallback);
    const double margin = waves >= 9L * cus ? 1.10 : 1.03;
    size_t fast = 0;
    for (size_t c = 1; c < cand.size(); ++c)
        if (best[c] < best[fast]) fast = c;
    int pick = cand[fast];
    double pick_ms = best[fast];
    for (size_t c = 0; c < cand.size(); ++c)
        if (cand[c] == fallback && best[c] <= margin * best[fast]) { pick = fallback; pick_ms = best[c]; }
    h->march_rows = pick;
    h->P.rows = pick;
    // the tiled kernels replace the march ones only by the same margin
    const bool tiled_wins = try_tiled && best_tiled * margin < pick_ms;
    if (tiled_wins) h->fused = KV_TILED;
    if (const char* dbg = std::getenv("SFV_TUNE_REPORT"); dbg && dbg[0] == '1') {
        std::fprintf(stderr, "[sfv autotune] %dx%d %s:", h->cfg.nx, h->cfg.ny, f64 ? "f64" : "f32");
        for (size_t c = 0; c < cand.size(); ++c) std::fprintf(stderr, " %d:%.4f", cand[c], best[c]);
        if (try_tiled) std::fprintf(stderr, " tiled:%.4f", best_tiled);
        std::fprintf(stderr, " -> %s%d (ms per step, %d steps per timing)\n", tiled_wins ? "tiled, march " : "",
                     pick, n);
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.emplace(key, tiled_wins ? -pick : pick);
    return SFV_OK;
}
}  // namespace sfv

namespace {

// update_bc! on the handle's current state (no step closing), on its stream.
int fill_ghosts(sfv_handle* h) {
    const int slot = (int)(h->step_parity & 1);
    if (h->cfg.precision == SFV_F64)
        HIP_TRY(h, launch_bc<double>(h->stream, h->P, h->ctl, slot, (double*)h->U[h->cur & 1], BC_STANDALONE));
    else
        HIP_TRY(h, launch_bc<float>(h->stream, h->P, h->ctl, slot, (float*)h->U[h->cur & 1], BC_STANDALONE));
    return SFV_OK;
}

// Run a prepared step loop of at most `nsteps` launched steps, in batches, and
// finish when the device says the loop ended.
int run_loop(sfv_handle* h, double t0, double endtime, int match_endtime, int has_max, double max_ts,
             int64_t max_steps, int until_endtime, int64_t* steps_out, double* t_out) {
    {
        std::string why;
        const int rt = sfv::tune_march_rows(h, why);
        if (rt) return fail(h, rt, why);
    }
    const int slot0 = (int)(h->step_parity & 1);
    HIP_TRY(h, launch_prepare(h->stream, h->P, h->ctl, slot0, t0, endtime, match_endtime, has_max, max_ts,
                              (long long)max_steps, until_endtime, h->cfg.precision));
    int64_t launched = 0;
    int64_t batch = until_endtime ? 8 : std::max<int64_t>(max_steps, 1);
    const int64_t cap = 256;
    for (;;) {
        int64_t nb = batch;
        if (max_steps > 0) nb = std::min<int64_t>(nb, max_steps - launched);
        if (nb <= 0) break;
        // steps launched past the end of the loop are no-ops on the device
        const int64_t save_cur = h->cur, save_par = h->step_parity;
        h->cur = (save_cur + launched) & 1;
        h->step_parity = (save_par + launched) & 1;
        const int rc = launch_steps(h, nb);
        h->cur = save_cur;
        h->step_parity = save_par;
        if (rc) return rc;
        launched += nb;
        if (!until_endtime) break;
        const int rc2 = collect_ctl(h);
        if (rc2) return rc2;
        if (!h->ctl_host->slot[(slot0 + launched) & 1].active) break;
        batch = std::min<int64_t>(batch * 2, cap);
    }
    const int rc = collect_ctl(h);
    if (rc) return rc;
    const int64_t done = h->ctl_host->steps;
    h->cur = (h->cur + done) & 1;
    h->step_parity = (h->step_parity + done) & 1;
    if (done > 0 && h->leaves_ghosts()) {
        // the fused kernels leave the output's ghost cells unwritten
        const int rf = fill_ghosts(h);
        if (rf) return rf;
    }
    if (done > 0) { h->last_dt = h->ctl_host->dt; h->have_dt = true; }
    if (steps_out) *steps_out = done;
    if (t_out) *t_out = h->ctl_host->t;
    if (h->ctl_host->nonfinite) return fail(h, SFV_ERR_NONFINITE_DT, "a time step produced a non-finite dt");
    return SFV_OK;
}

}  // namespace

namespace sfv {
int create_handle(const sfv_config* cfg, DevicePool* pool, sfv_handle** out, std::string& why) {
    *out = nullptr;
    const int vr = validate_config(cfg, why);
    if (vr) return vr;
    sfv_handle* h = new sfv_handle();
    h->pool = pool;
    const int rc = h->init(*cfg, why);
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return SFV_OK;
}
}  // namespace sfv

extern "C" {

void sfv_config_defaults(sfv_config* c, int32_t nx, int32_t ny, double dx, double dy) {
    std::memset(c, 0, sizeof(*c));
    c->nx = nx; c->ny = ny; c->gc = 2;
    c->dx = dx; c->dy = dy;
    c->bc = SFV_BC_WALL;
    c->g = 9.81; c->rho = 1.0; c->depth_cutoff = 1e-5; c->desingularizing_kappa = 1e-5;
    c->reconstruction = SFV_LINEAR_RECONSTRUCTION; c->theta = 1.2;
    c->stepper = SFV_RUNGE_KUTTA2; c->cfl = 0.25;
    c->Cz = 0.03 * 0.03;
    c->horton_f0 = 1.977e-4; c->horton_fc = 3.272e-5; c->horton_k = 2.43e-3;
    c->infiltration_factor_scalar = 1.0;
    c->rain_before_infiltration = 1;
    c->precision = SFV_F64;
}

const char* sfv_create_error(void) { return g_create_error.c_str(); }

const char* sfv_last_error(const sfv_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int64_t sfv_device_bytes_for(const sfv_config* c) {
    if (!c) return -1;
    const int64_t nxT = c->nx + 2 * (int64_t)c->gc, nyT = c->ny + 2 * (int64_t)c->gc;
    const int64_t elem = c->precision == SFV_F32_STORAGE ? 4 : 8;
    // ping-pong pair + right-hand side (+ the stage-1 state of the per-pass
    // path, which RK2 grids the fused kernels cannot take use)
    int64_t nstate = 3;
    if (c->stepper == SFV_RUNGE_KUTTA2) {
        sfv_handle probe;
        probe.cfg = *c;
        probe.build_params();
        if (!probe.fusable()) nstate = 4;
    }
    int64_t b = nstate * 3 * nxT * nyT * elem + (nxT + 1) * (nyT + 1) * 8 + 4096;
    if (c->infiltration == SFV_INFILTRATION_HORTON) b += nxT * nyT * 8;
    return b;
}

int sfv_create(const sfv_config* cfg, sfv_handle** out) {
    if (!out) return fail(nullptr, SFV_ERR_ARGUMENT, "out is NULL");
    std::string why;
    const int rc = sfv::create_handle(cfg, nullptr, out, why);
    if (rc) return fail(nullptr, rc, why);
    return SFV_OK;
}

void sfv_destroy(sfv_handle* h) { delete h; }

int sfv_set_topography(sfv_handle* h, const double* B, int64_t nbx, int64_t nby) {
    if (!h || !B) return fail(h, SFV_ERR_ARGUMENT, "NULL argument");
    const KParams& P = h->P;
    if (nbx != P.nxT + 1 || nby != P.nyT + 1) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "Bottom topography should be of size (%d, %d) but got (%lld, %lld)", P.nxT + 1,
                      P.nyT + 1, (long long)nbx, (long long)nby);
        return fail(h, SFV_ERR_DOMAIN, buf);
    }
    const size_t nb = (size_t)nbx * nby;
    bool nonzero = false;
    for (size_t k = 0; k < nb && !nonzero; ++k) nonzero = !(B[k] == 0.0);
    if (nonzero && !h->cfg.bottom_source)
        return fail(h, SFV_ERR_ARGUMENT,
                    "Found non-zero bottom topography in equation, but no corresponding source term. "
                    "Did you forget to add a source term to your system?");
    HIP_TRY(h, hipSetDevice(h->device));
    HIP_TRY(h, hipMemcpyAsync(h->B, B, nb * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return SFV_OK;
}

int sfv_set_rain_field(sfv_handle* h, const double* field) {
    if (!h || !field) return fail(h, SFV_ERR_ARGUMENT, "NULL argument");
    if (h->cfg.rain != SFV_RAIN_FIELD || !h->d_rain_field)
        return fail(h, SFV_ERR_ARGUMENT, "the handle was not created with rain = SFV_RAIN_FIELD");
    HIP_TRY(h, hipSetDevice(h->device));
    HIP_TRY(h, hipMemcpyAsync(h->d_rain_field, field, h->ncell * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return SFV_OK;
}

int sfv_set_state(sfv_handle* h, const void* U, int32_t layout) {
    if (!h || !U) return fail(h, SFV_ERR_ARGUMENT, "NULL argument");
    HIP_TRY(h, hipSetDevice(h->device));
    const KParams& P = h->P;
    const size_t es = h->elem;
    void* dst = h->U[h->cur];
    if (layout == SFV_LAYOUT_FULL) {
        HIP_TRY(h, hipMemcpyAsync(dst, U, 3 * h->ncell * es, hipMemcpyHostToDevice, h->stream));
    } else if (layout == SFV_LAYOUT_INTERIOR) {
        // one contiguous upload into the right-hand-side buffer (dead between
        // steps), then a device scatter into the padded state
        HIP_TRY(h, hipMemcpyAsync(h->RHS, U, 3 * (size_t)P.nx * P.ny * es, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(h, launch_interior_copy(h->stream, P, h->RHS, dst, es, 1));
        const int rc = h->cfg.precision == SFV_F64
                           ? (int)launch_bc<double>(h->stream, P, h->ctl, 0, (double*)dst, BC_STANDALONE)
                           : (int)launch_bc<float>(h->stream, P, h->ctl, 0, (float*)dst, BC_STANDALONE);
        if (rc) return fail(h, SFV_ERR_HIP, "boundary-condition launch failed");
    } else {
        return fail(h, SFV_ERR_ARGUMENT, "unknown layout");
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    h->has_state = true;
    return SFV_OK;
}

int sfv_get_state(sfv_handle* h, void* U, int32_t layout) {
    if (!h || !U) return fail(h, SFV_ERR_ARGUMENT, "NULL argument");
    HIP_TRY(h, hipSetDevice(h->device));
    const KParams& P = h->P;
    const size_t es = h->elem;
    const void* src = h->U[h->cur];
    if (layout == SFV_LAYOUT_FULL) {
        HIP_TRY(h, hipMemcpyAsync(U, src, 3 * h->ncell * es, hipMemcpyDeviceToHost, h->stream));
    } else if (layout == SFV_LAYOUT_INTERIOR) {
        // device gather into the right-hand-side buffer, one contiguous download
        HIP_TRY(h, launch_interior_copy(h->stream, P, h->RHS, const_cast<void*>(src), es, 0));
        HIP_TRY(h, hipMemcpyAsync(U, h->RHS, 3 * (size_t)P.nx * P.ny * es, hipMemcpyDeviceToHost, h->stream));
    } else {
        return fail(h, SFV_ERR_ARGUMENT, "unknown layout");
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return SFV_OK;
}

int sfv_step(sfv_handle* h, double max_dt, double t, double* dt_out) {
    int rc = ensure_ready(h);
    if (rc) return rc;
    HIP_TRY(h, hipSetDevice(h->device));
    int64_t steps = 0;
    double tend = 0;
    rc = run_loop(h, t, NAN, 0, 1, max_dt, 1, 0, &steps, &tend);
    if (dt_out) *dt_out = h->last_dt;
    return rc;
}

int sfv_run(sfv_handle* h, double t_end, int32_t match_endtime, double maximum_timestep, int64_t max_steps,
            double* t_io, int64_t* steps_out) {
    int rc = ensure_ready(h);
    if (rc) return rc;
    if (!t_io) return fail(h, SFV_ERR_ARGUMENT, "t_io is NULL");
    if (h->cfg.rain == SFV_RAIN_FIELD)
        return fail(h, SFV_ERR_ARGUMENT, "FunctionalRain needs its rate field set before every step: use sfv_step");
    HIP_TRY(h, hipSetDevice(h->device));
    const int has_max = maximum_timestep > 0 ? 1 : 0;
    return run_loop(h, *t_io, t_end, match_endtime ? 1 : 0, has_max, has_max ? maximum_timestep : 0.0,
                    max_steps, 1, steps_out, t_io);
}

int sfv_advance(sfv_handle* h, int64_t nsteps, double max_dt, double* t_io) {
    int rc = ensure_ready(h);
    if (rc) return rc;
    if (!t_io) return fail(h, SFV_ERR_ARGUMENT, "t_io is NULL");
    if (h->cfg.rain == SFV_RAIN_FIELD)
        return fail(h, SFV_ERR_ARGUMENT, "FunctionalRain needs its rate field set before every step: use sfv_step");
    if (nsteps <= 0) return SFV_OK;
    HIP_TRY(h, hipSetDevice(h->device));
    int64_t steps = 0;
    return run_loop(h, *t_io, NAN, 0, 1, max_dt, nsteps, 0, &steps, t_io);
}

int sfv_current_timestep(sfv_handle* h, double* dt_out) {
    if (!h || !dt_out) return fail(h, SFV_ERR_ARGUMENT, "NULL argument");
    *dt_out = h->have_dt ? h->last_dt : 0.0;
    return SFV_OK;
}

int sfv_synchronize(sfv_handle* h) {
    if (!h) return SFV_ERR_ARGUMENT;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return SFV_OK;
}

int64_t sfv_device_bytes(const sfv_handle* h) { return h ? h->device_bytes : 0; }

int sfv_time_kernel(sfv_handle* h, const char* which, int32_t n, double* ms_out) {
    int rc = ensure_ready(h);
    if (rc) return rc;
    if (!which || !ms_out || n < 1) return fail(h, SFV_ERR_ARGUMENT, "bad arguments");
    if (h->cfg.stepper != SFV_RUNGE_KUTTA2 || !h->fusable())
        return fail(h, SFV_ERR_ARGUMENT, "kernel timing needs RungeKutta2 with fused stage kernels");
    const std::string w(which);
    if (w != "stage1" && w != "stage2" && w != "bc" && w != "step" && w != "copy")
        return fail(h, SFV_ERR_ARGUMENT, "which must be stage1, stage2, bc, step or copy");
    HIP_TRY(h, hipSetDevice(h->device));
    {
        std::string why;
        int rc2 = sfv::tune_march_rows(h, why);
        if (!rc2) rc2 = h->ensure_scratch(why);
        if (rc2) return fail(h, rc2, why);
    }
    const int slot = (int)(h->step_parity & 1), cur = (int)(h->cur & 1);
    // The kernels run as in a step — stage 1, stage 2 (with the dt stage 1's
    // wave speeds give; both read ghosts through the boundary maps, so no
    // boundary kernel runs per step and "bc" times 0) — against a
    // scratch output without closing the step, so the state is left as is,
    // with an event between consecutive kernels; the named kernel's mean
    // duration is returned.  The first step is a warm-up.
    HIP_TRY(h, launch_prepare(h->stream, h->P, h->ctl, slot, 0.0, NAN, 0, 0, 0.0, -1, 0, h->cfg.precision));
    auto stage = [&](int which, auto* U0, auto* R, auto* Us) -> hipError_t {
        const KParams& P = h->P;
        if (h->fused == KV_TILED)
            return which == 1 ? launch_stage1(h->stream, P, h->ctl, slot, U0, R, h->B)
                              : launch_stage2(h->stream, P, h->ctl, slot, U0, R, Us, h->B);
        return which == 1 ? launch_march(h->stream, P, h->ctl, slot, 1, U0, (decltype(U0))nullptr, R, h->B, 0, 1)
                          : launch_march(h->stream, P, h->ctl, slot, 2, U0, R, Us, h->B);
    };
    auto kernel = [&](int k) -> hipError_t {      // 0 stage 1, 1 stage 2, 2 boundary condition, 3 copy
        if (k == 3) return launch_copy(h->stream, h->U[cur], h->U[2], 3 * h->ncell * h->elem);
        if (h->cfg.precision == SFV_F64) {
            double *U0 = (double*)h->U[cur], *R = (double*)h->RHS, *Us = (double*)h->U[2];
            return k == 2 ? launch_bc<double>(h->stream, h->P, h->ctl, slot, Us, BC_IN_STEP) : stage(k + 1, U0, R, Us);
        }
        float *U0 = (float*)h->U[cur], *R = (float*)h->RHS, *Us = (float*)h->U[2];
        return k == 2 ? launch_bc<float>(h->stream, h->P, h->ctl, slot, Us, BC_IN_STEP) : stage(k + 1, U0, R, Us);
    };
    const bool copy = w == "copy";
    const int nk = copy ? 1 : 2;
    const int runs = n + 1;
    std::vector<hipEvent_t> ev((size_t)runs * (nk + 1));
    for (auto& e : ev) HIP_TRY(h, hipEventCreate(&e));
    hipError_t le = hipSuccess;
    for (int r = 0; r < runs && le == hipSuccess; ++r)
        for (int k = 0; k <= nk && le == hipSuccess; ++k) {
            le = hipEventRecord(ev[(size_t)r * (nk + 1) + k], h->stream);
            if (k < nk && le == hipSuccess) le = kernel(copy ? 3 : k);
        }
    if (le == hipSuccess) le = hipEventSynchronize(ev.back());
    double tot[4] = {0, 0, 0, 0};
    for (int r = 1; r < runs && le == hipSuccess; ++r)
        for (int k = 0; k < nk && le == hipSuccess; ++k) {
            float ms = 0;
            le = hipEventElapsedTime(&ms, ev[(size_t)r * (nk + 1) + k], ev[(size_t)r * (nk + 1) + k + 1]);
            tot[k] += ms;
        }
    for (auto& e : ev) (void)hipEventDestroy(e);
    HIP_TRY(h, le);
    double ms = 0;
    if (copy || w == "stage1") ms = tot[0];
    else if (w == "stage2") ms = tot[1];
    else if (w == "bc") ms = tot[2];
    else ms = tot[0] + tot[1] + tot[2];
    *ms_out = ms / n;
    return SFV_OK;
}

int sfv_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int sfv_set_fused(sfv_handle* h, int32_t fused) {
    if (!h) return SFV_ERR_ARGUMENT;
    if (fused < KV_PER_PASS || fused > KV_HYBRID) return fail(h, SFV_ERR_ARGUMENT, "unknown kernel variant");
    if (fused == KV_HYBRID && !h->hybrid_ok())
        return fail(h, SFV_ERR_ARGUMENT, "the march stage-1 kernel needs gc >= 2");
    if ((fused == KV_TILED || fused == KV_MARCH) && !h->fusable())
        return fail(h, SFV_ERR_ARGUMENT,
                    "fused kernels need RungeKutta2, gc >= 2 and boundary-condition sources that are interior cells");
    h->fused = fused;
    h->variant_forced = true;
    return SFV_OK;
}

int sfv_set_march_rows(sfv_handle* h, int32_t rows) {
    if (!h) return SFV_ERR_ARGUMENT;
    if (rows < 0 || rows > 1024) return fail(h, SFV_ERR_ARGUMENT, "rows must be in 0..1024");
    h->march_rows = rows;
    h->rows_forced = rows > 0;
    h->P.rows = rows;
    return SFV_OK;
}

int32_t sfv_get_march_rows(const sfv_handle* h) { return h ? h->march_rows : -1; }

int32_t sfv_get_fused(const sfv_handle* h) { return h ? h->fused : -1; }

}  // extern "C"
