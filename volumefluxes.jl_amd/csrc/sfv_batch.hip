// sfv_batch.hip — ensembles: several handles (grids of any size, each with its
// own configuration, state, time and time step) advanced together, one
// persistent launch per RK stage for the whole batch, the step sequence
// captured once in a HIP graph and replayed.
//
// Each member keeps its own buffers; the batch owns the member table the
// kernels read, the members' step-control blocks (one contiguous array, so a
// single copy reads every member's progress) and the graph.  Results are
// bitwise those of stepping each member alone: the kernels are the same
// per-wave tile code, only the block-to-tile map differs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sinfvm_hip.h"
#include "sfv_engine.hpp"

using namespace sfv;

namespace sfv {

BatchEngine::~BatchEngine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_) (void)hipGraphExecDestroy(graph_);
    for (hipEvent_t e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (d_mem_) (void)hipFree(d_mem_);
    if (d_hdr_) (void)hipFree(d_hdr_);
    if (d_ctl_) (void)hipFree(d_ctl_);
    if (h_mem_) (void)hipHostFree(h_mem_);
    if (h_hdr_) (void)hipHostFree(h_hdr_);
    if (h_ctl_) (void)hipHostFree(h_ctl_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

#define ETRY(expr)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) {                                                                      \
            why = std::string(#expr) + ": " + hipGetErrorString(e_);                                 \
            return SFV_ERR_HIP;                                                                      \
        }                                                                                            \
    } while (0)

int BatchEngine::init(int precision, int capacity, std::string& why) {
    precision_ = precision;
    ETRY(hipGetDevice(&device_));
    ETRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    ETRY(hipEventCreate(&ev_[0]));
    ETRY(hipEventCreate(&ev_[1]));
    ETRY(hipMalloc((void**)&d_hdr_, sizeof(BatchHeader)));
    ETRY(hipHostMalloc((void**)&h_hdr_, sizeof(BatchHeader), hipHostMallocDefault));
    hipDeviceProp_t prop;
    int cus = 256;
    if (hipGetDeviceProperties(&prop, device_) == hipSuccess && prop.multiProcessorCount > 0)
        cus = prop.multiProcessorCount;
    cus_ = cus;
    per_cu_ = std::min(march_batch_blocks_per_cu(precision, 1), march_batch_blocks_per_cu(precision, 2));
    size_grid();
    return grow(std::max(1, capacity), why);
}

void BatchEngine::size_grid() {
    const int per = share_ == Share::ALONE ? per_cu_ : share_ == Share::HOST ? std::max(1, per_cu_ - 1) : 1;
    blocks_stage_ = std::max(8, (cus_ * per + 7) / 8 * 8);
}

void BatchEngine::set_share(Share s) {
    if (s == share_) return;
    share_ = s;
    size_grid();
    if (graph_) {
        (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    table_dirty_ = true;   // segment heights and the priority flag follow the grid
}

bool BatchEngine::replay_ready() const { return hipStreamQuery(stream_) != hipErrorNotReady; }

int BatchEngine::grow(int capacity, std::string& why) {
    BatchMember* dm = nullptr;
    StepCtl* dc = nullptr;
    BatchMember* hm = nullptr;
    StepCtl* hc = nullptr;
    if (hipMalloc((void**)&dm, sizeof(BatchMember) * capacity) != hipSuccess ||
        hipMalloc((void**)&dc, sizeof(StepCtl) * capacity) != hipSuccess ||
        hipHostMalloc((void**)&hm, sizeof(BatchMember) * capacity, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&hc, sizeof(StepCtl) * capacity, hipHostMallocDefault) != hipSuccess) {
        if (dm) (void)hipFree(dm);
        if (dc) (void)hipFree(dc);
        if (hm) (void)hipHostFree(hm);
        if (hc) (void)hipHostFree(hc);
        why = "allocation of the batch tables failed";
        return SFV_ERR_OUT_OF_MEMORY;
    }
    std::memset(hc, 0, sizeof(StepCtl) * capacity);
    if (d_ctl_) {   // members keep their step control across the move
        ETRY(hipMemcpyAsync(dc, d_ctl_, sizeof(StepCtl) * capacity_, hipMemcpyDeviceToDevice, stream_));
        std::memcpy(hc, h_ctl_, sizeof(StepCtl) * capacity_);
        ETRY(hipStreamSynchronize(stream_));
        (void)hipFree(d_mem_);
        (void)hipFree(d_ctl_);
        (void)hipHostFree(h_mem_);
        (void)hipHostFree(h_ctl_);
    }
    d_mem_ = dm; d_ctl_ = dc; h_mem_ = hm; h_ctl_ = hc;
    capacity_ = capacity;
    slot_handle_.resize(capacity, nullptr);
    slot_row_.resize(capacity);
    ran_.resize(capacity, 0);
    for (int s = 0; s < capacity; ++s) slot_row_[s].ctl = d_ctl_ + s;
    if (graph_) {   // the graph holds the old table pointer
        (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    table_dirty_ = true;
    return SFV_OK;
}

int BatchEngine::add(sfv_handle* h, const RunParams& rp, int* slot_out, std::string& why) {
    if (h->cfg.stepper != SFV_RUNGE_KUTTA2 || !h->fusable()) {
        why = "batched members run the fused RK2 kernels (RungeKutta2, ghost sources inside the grid)";
        return SFV_ERR_ARGUMENT;
    }
    if (h->cfg.precision != precision_) { why = "all members must have the same precision"; return SFV_ERR_ARGUMENT; }
    if (h->cfg.rain == SFV_RAIN_FIELD) {
        why = "FunctionalRain needs its rate field set before every step; it cannot run batched";
        return SFV_ERR_ARGUMENT;
    }
    if (h->device != device_) { why = "all members must live on one device"; return SFV_ERR_ARGUMENT; }
    int s = -1;
    for (int k = 0; k < capacity_ && s < 0; ++k)
        if (!slot_handle_[k]) s = k;
    if (s < 0) {
        s = capacity_;
        const int rc = grow(2 * capacity_, why);
        if (rc) return rc;
    }
    ETRY(hipStreamSynchronize(h->stream));   // pending set_state / topography uploads
    BatchMember M;
    std::memset(&M, 0, sizeof(M));
    M.P = h->P;
    M.ctl = d_ctl_ + s;
    M.U[0] = h->U[0];
    M.U[1] = h->U[1];
    M.RHS = h->RHS;
    M.B = h->B;
    M.cur0 = (int)(h->cur & 1);
    M.slot0 = (int)(h->step_parity & 1);
    M.my = h->P.rows > 0 ? h->P.rows : march_rows_per_wave(h->P, 1, h->cfg.precision == SFV_F64 ? 0 : 1);
    const int segs = (h->P.ny + M.my - 1) / M.my;
    M.strips = (h->P.nx + MARCH_COLS - 1) / MARCH_COLS;
    M.segs = segs;
    M.ntiles = march_blocks(M.strips, segs);
    M.t0 = rp.t0;
    M.until_endtime = rp.until_endtime;
    M.endtime = rp.until_endtime ? rp.endtime : NAN;
    M.match_endtime = rp.match_endtime;
    M.has_max_timestep = rp.maximum_timestep > 0 ? 1 : 0;
    M.max_timestep = rp.maximum_timestep > 0 ? rp.maximum_timestep : 0.0;
    M.max_steps = rp.max_steps;
    M.precision = precision_;
    ETRY(launch_prepare(stream_, h->P, M.ctl, M.slot0, M.t0, M.endtime, M.match_endtime, M.has_max_timestep,
                        M.max_timestep, M.max_steps, M.until_endtime, precision_));
    slot_row_[s] = M;
    slot_handle_[s] = h;
    ran_[s] = 0;
    rows_.push_back(s);
    table_dirty_ = true;
    members_changed_ = true;
    *slot_out = s;
    return SFV_OK;
}

void BatchEngine::remove(int slot) {
    sfv_handle* h = slot_handle_[slot];
    if (!h) return;
    if (ran_[slot]) {
        const StepCtl& c = h_ctl_[slot];
        h->cur = (h->cur + c.steps) & 1;
        h->step_parity = (h->step_parity + c.steps) & 1;
        if (c.steps > 0) {
            h->last_dt = c.dt;
            h->have_dt = true;
            // the march kernels leave the output's ghost cells unwritten: fill
            // them on the member's own stream before it is read or stepped alone
            const int ps = (int)(h->step_parity & 1);
            if (precision_ == SFV_F64)
                (void)launch_bc<double>(h->stream, h->P, h->ctl, ps, (double*)h->U[h->cur & 1], BC_STANDALONE);
            else
                (void)launch_bc<float>(h->stream, h->P, h->ctl, ps, (float*)h->U[h->cur & 1], BC_STANDALONE);
        }
    }
    slot_handle_[slot] = nullptr;
    ran_[slot] = 0;
    for (size_t r = 0; r < rows_.size(); ++r)
        if (rows_[r] == slot) { rows_.erase(rows_.begin() + (long)r); break; }
    table_dirty_ = true;
    members_changed_ = true;
}

bool BatchEngine::finished(int slot) const {
    if (!ran_[slot]) return false;
    // replays end on launch parity 0, where the member's next slot is slot0
    return h_ctl_[slot].slot[slot_row_[slot].slot0].active == 0;
}

// Segment heights of the members for the current membership.  When the
// batch fills the device's resident wave slots at the batch default height
// (SFV_BATCH_DEFAULT_ROWS),
// every member takes it (fewest halo instructions).  Otherwise the device is
// underfilled and a step lasts about one wave's serial path, which shorter
// segments shorten as long as no SIMD has to run two of the batch's waves:
// every member takes the shortest height at which the batch's waves fit one
// per SIMD (the default if none does).  Heights set with sfv_set_march_rows
// are kept.  No height changes any result.  The online tuner (set_autotune)
// times this choice against its neighbours.
int BatchEngine::rule_rows() const {
    const int R0 = SFV_BATCH_DEFAULT_ROWS;
    const long slots = (long)blocks_stage_ * MARCH_WAVES, simds = 4L * cus_;
    auto waves = [&](int r) {
        long w = 0;
        for (int s : rows_) {
            const sfv_handle* h = slot_handle_[s];
            const int rr = h->rows_forced ? h->march_rows : std::min(r, h->P.ny);
            w += (long)((h->P.nx + MARCH_COLS - 1) / MARCH_COLS) * ((h->P.ny + rr - 1) / rr);
        }
        return w;
    };
    int r = R0;
    if (waves(R0) < slots)
        for (int c : {1, 2, 3, 4, 6, 8})
            if (waves(c) <= simds) { r = c; break; }
    return r;
}

void BatchEngine::choose_rows() {
    const int r = rows_override_ > 0 ? rows_override_ : rule_rows();
    rows_in_use_ = rows_.empty() ? 0 : r;
    for (int s : rows_) {
        const sfv_handle* h = slot_handle_[s];
        BatchMember& M = slot_row_[s];
        const int my = h->rows_forced ? h->march_rows : r;
        M.my = my;
        M.P.rows = my;
        const int segs = (h->P.ny + my - 1) / my;
        M.segs = segs;
        M.ntiles = march_blocks(M.strips, segs);
    }
}

// What a kept height choice depends on: the share of the device and every
// member's size and forced height.
std::vector<long long> BatchEngine::signature() const {
    std::vector<long long> sig;
    sig.push_back(blocks_stage_);
    for (int s : rows_) {
        const sfv_handle* h = slot_handle_[s];
        sig.push_back(((long long)h->P.nx << 32) | (unsigned)h->P.ny);
        sig.push_back(h->rows_forced ? h->march_rows : 0);
    }
    std::sort(sig.begin() + 1, sig.end());
    return sig;
}

// After a membership change: reuse the kept choice for the same member sizes,
// else line up the timed trials: the rule's height, the next lower and the
// three next higher candidates, the rule's height again.
void BatchEngine::plan_tuning() {
    trials_.clear();
    trial_ms_.clear();
    trial_ = 0;
    rows_override_ = 0;
    if (!autotune_ || rows_.empty()) return;
    bool any_free = false;
    for (int s : rows_) any_free = any_free || !slot_handle_[s]->rows_forced;
    if (!any_free) return;
    const std::vector<long long> sig = signature();
    if (sig == tuned_sig_) { rows_override_ = tuned_rows_; return; }
    static const int C[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48};
    const int n = (int)(sizeof(C) / sizeof(C[0]));
    const int r0 = rule_rows();
    int i0 = 0;
    while (i0 < n - 1 && C[i0] < r0) ++i0;
    trials_.push_back(r0);
    if (i0 > 0) trials_.push_back(C[i0 - 1]);
    if (i0 < n - 1) trials_.push_back(C[i0 + 1]);
    if (i0 < n - 2) trials_.push_back(C[i0 + 2]);
    if (i0 < n - 3) trials_.push_back(C[i0 + 3]);
    trials_.push_back(r0);
    trial_ms_.assign(trials_.size(), 0.f);
    rows_override_ = trials_[0];
}

int BatchEngine::upload_table(std::string& why) {
    choose_rows();
    int tiles = 0;
    for (size_t r = 0; r < rows_.size(); ++r) {
        BatchMember& M = slot_row_[rows_[r]];
        M.tile0 = tiles;
        tiles += M.ntiles;
        h_mem_[r] = M;
    }
    h_hdr_->members = (int)rows_.size();
    h_hdr_->tiles = tiles;
    h_hdr_->prio = share_ == Share::GUEST ? 1 : share_ == Share::HOST ? 2 : 0;
    h_hdr_->pad = 0;
    if (!rows_.empty())
        ETRY(hipMemcpyAsync(d_mem_, h_mem_, sizeof(BatchMember) * rows_.size(), hipMemcpyHostToDevice, stream_));
    ETRY(hipMemcpyAsync(d_hdr_, h_hdr_, sizeof(BatchHeader), hipMemcpyHostToDevice, stream_));
    table_dirty_ = false;
    return SFV_OK;
}

// graph_steps RK2 steps (stage 1, then stage 2, which also fills the ghosts
// and closes each member's step), launch parity alternating 0, 1, ... on the
// engine's stream.
int BatchEngine::capture(std::string& why) {
    ETRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int s = 0; s < graph_steps_ && e == hipSuccess; ++s) {
        const int lp = s & 1;
        if (precision_ == SFV_F64) {
            e = launch_march_batch<double>(stream_, 1, d_mem_, d_hdr_, lp, blocks_stage_, fast_);
            if (e == hipSuccess) e = launch_march_batch<double>(stream_, 2, d_mem_, d_hdr_, lp, blocks_stage_, fast_);
        } else {
            e = launch_march_batch<float>(stream_, 1, d_mem_, d_hdr_, lp, blocks_stage_, fast_);
            if (e == hipSuccess) e = launch_march_batch<float>(stream_, 2, d_mem_, d_hdr_, lp, blocks_stage_, fast_);
        }
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(stream_, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        why = std::string("batched launch during capture: ") + hipGetErrorString(e);
        return SFV_ERR_HIP;
    }
    if (e2 != hipSuccess) { why = std::string("hipStreamEndCapture: ") + hipGetErrorString(e2); return SFV_ERR_HIP; }
    const hipError_t e3 = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) {
        graph_ = nullptr;
        why = std::string("hipGraphInstantiate: ") + hipGetErrorString(e3);
        return SFV_ERR_HIP;
    }
    return SFV_OK;
}

int BatchEngine::replay_begin(int k, std::string& why) {
    int rc;
    if (table_dirty_) {
        // the graph holds the specialised kernels only while every member fits them
        bool fast = !rows_.empty();
        for (int s : rows_) fast = fast && march_fast(slot_handle_[s]->P);
        if (graph_ && fast != fast_) {
            (void)hipGraphExecDestroy(graph_);
            graph_ = nullptr;
        }
        fast_ = fast;
        if (members_changed_) {
            members_changed_ = false;
            plan_tuning();
        }
        if ((rc = upload_table(why))) return rc;
    }
    if (!graph_ && (rc = capture(why))) return rc;
    timed_k_ = tuning() ? k : 0;
    if (timed_k_) ETRY(hipEventRecord(ev_[0], stream_));
    for (int r = 0; r < k; ++r) ETRY(hipGraphLaunch(graph_, stream_));
    if (timed_k_) ETRY(hipEventRecord(ev_[1], stream_));
    ETRY(hipMemcpyAsync(h_ctl_, d_ctl_, sizeof(StepCtl) * capacity_, hipMemcpyDeviceToHost, stream_));
    return SFV_OK;
}

int BatchEngine::replay_end(std::string& why) {
    ETRY(hipStreamSynchronize(stream_));
    for (int s : rows_) ran_[s] = 1;      // membership cannot change between begin and end
    if (timed_k_) {
        float ms = 0.f;
        ETRY(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
        trial_ms_[trial_] = ms / (float)timed_k_;
        timed_k_ = 0;
        if (++trial_ < (int)trials_.size()) {
            rows_override_ = trials_[trial_];
        } else {
            // A step's cost drifts as the flow develops, so each trial is set
            // against the rule's height timed first and last, interpolated to
            // the trial's place; another height must be 3% faster than that.
            // Heights within 1.5% of the best count as tied, and the one
            // nearest the rule's height (in ratio) wins the tie.
            const int last = trial_ - 1, r0 = trials_[last];
            std::vector<float> ratio(trials_.size(), 1.f);
            float best_ratio = 0.97f;
            for (int i = 1; i < last; ++i) {
                const float ref = trial_ms_[0] + (trial_ms_[last] - trial_ms_[0]) * (float)i / (float)last;
                ratio[i] = ref > 0.f ? trial_ms_[i] / ref : 1.f;
                best_ratio = std::min(best_ratio, ratio[i]);
            }
            int best = r0;
            double best_dist = 1e30;
            for (int i = 1; i < last; ++i) {
                const double dist = std::fabs(std::log((double)trials_[i] / r0));
                if (ratio[i] < 0.97f && ratio[i] <= best_ratio + 0.015f && dist < best_dist) {
                    best = trials_[i];
                    best_dist = dist;
                }
            }
            rows_override_ = best;
            tuned_rows_ = best;
            tuned_sig_ = signature();
            if (const char* e = std::getenv("SFV_TUNE_REPORT"); e && e[0] == '1') {
                std::fprintf(stderr, "[sfv batch autotune] %d members:", members());
                for (size_t i = 0; i < trials_.size(); ++i) std::fprintf(stderr, " %d:%.3fms", trials_[i], trial_ms_[i]);
                std::fprintf(stderr, " -> %d\n", best);
            }
        }
        if (rows_override_ != rows_in_use_) table_dirty_ = true;
    }
    return SFV_OK;
}

int BatchEngine::replay(int k, std::string& why) {
    const int rc = replay_begin(k, why);
    if (rc) return rc;
    return replay_end(why);
}

#undef ETRY

}  // namespace sfv

struct sfv_batch {
    std::vector<sfv_handle*> members;
    sfv::BatchEngine engine;
    std::string err;
};

namespace {

thread_local std::string g_batch_error;

int bfail(sfv_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    else g_batch_error = msg;
    return code;
}

}  // namespace

extern "C" {

int sfv_batch_create(sfv_handle* const* handles, int32_t n, sfv_batch** out) {
    if (!out) return bfail(nullptr, SFV_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!handles || n < 1) return bfail(nullptr, SFV_ERR_ARGUMENT, "a batch needs at least one member");
    for (int i = 0; i < n; ++i) {
        sfv_handle* h = handles[i];
        if (!h) return bfail(nullptr, SFV_ERR_ARGUMENT, "member " + std::to_string(i) + " is NULL");
        for (int k = 0; k < i; ++k)
            if (handles[k] == h) return bfail(nullptr, SFV_ERR_ARGUMENT, "a handle appears twice in the batch");
        if (h->cfg.stepper != SFV_RUNGE_KUTTA2 || !h->fusable())
            return bfail(nullptr, SFV_ERR_ARGUMENT,
                         "member " + std::to_string(i) +
                             ": batches run the fused RK2 kernels (RungeKutta2, ghost sources inside the grid)");
        if (h->cfg.rain == SFV_RAIN_FIELD)
            return bfail(nullptr, SFV_ERR_ARGUMENT,
                         "member " + std::to_string(i) + ": FunctionalRain needs its rate field every step");
        if (h->device != handles[0]->device)
            return bfail(nullptr, SFV_ERR_ARGUMENT, "all members must live on one device");
        if (h->cfg.precision != handles[0]->cfg.precision)
            return bfail(nullptr, SFV_ERR_ARGUMENT, "all members must have the same precision");
    }
    sfv_batch* b = new sfv_batch();
    b->members.assign(handles, handles + n);
    std::string why;
    int rc = hipSetDevice(handles[0]->device) == hipSuccess ? SFV_OK : SFV_ERR_HIP;
    if (rc == SFV_OK) rc = b->engine.init(handles[0]->cfg.precision, n, why);
    else why = "hipSetDevice failed";
    if (rc) {
        delete b;
        return bfail(nullptr, rc, why);
    }
    b->engine.set_autotune(sfv::autotune_enabled());
    *out = b;
    return SFV_OK;
}

void sfv_batch_destroy(sfv_batch* b) { delete b; }

const char* sfv_batch_last_error(const sfv_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

int32_t sfv_batch_size(const sfv_batch* b) { return b ? (int32_t)b->members.size() : 0; }

int sfv_batch_get_march_rows(const sfv_batch* b, int32_t* rows) {
    if (!b || !rows) return SFV_ERR_ARGUMENT;
    *rows = b->engine.rows_in_use();
    return SFV_OK;
}

int sfv_batch_run(sfv_batch* b, const double* t_end, int32_t match_endtime, double maximum_timestep,
                  int64_t max_steps, double* t_io, int64_t* steps_out) {
    if (!b) return SFV_ERR_ARGUMENT;
    if (!t_io) return bfail(b, SFV_ERR_ARGUMENT, "t_io is NULL");
    if (!t_end && max_steps <= 0) return bfail(b, SFV_ERR_ARGUMENT, "give t_end or a positive max_steps");
    const int n = (int)b->members.size();
    for (int i = 0; i < n; ++i)
        if (!b->members[i]->has_state) return bfail(b, SFV_ERR_STATE, "member " + std::to_string(i) + " has no state");
    if (hipSetDevice(b->members[0]->device) != hipSuccess) return bfail(b, SFV_ERR_HIP, "hipSetDevice failed");
    std::string why;
    std::vector<int> slots(n, -1);
    auto drop_all = [&]() {
        for (int i = 0; i < n; ++i)
            if (slots[i] >= 0) b->engine.remove(slots[i]);
    };
    for (int i = 0; i < n; ++i) {
        sfv::RunParams rp;
        rp.t0 = t_io[i];
        rp.until_endtime = t_end ? 1 : 0;
        rp.endtime = t_end ? t_end[i] : NAN;
        rp.match_endtime = match_endtime ? 1 : 0;
        rp.maximum_timestep = maximum_timestep;
        rp.max_steps = max_steps;
        const int rc = b->engine.add(b->members[i], rp, &slots[i], why);
        if (rc) { drop_all(); return bfail(b, rc, why); }
    }
    // Replay until every member has finished; a fixed step count needs no
    // look at the device on the way.
    const int gs = b->engine.graph_steps();
    int64_t launched = 0;
    int per_check = 1;
    for (;;) {
        int64_t k = per_check;
        if (max_steps > 0) k = std::min<int64_t>(k, (max_steps - launched + gs - 1) / gs);
        if (!t_end) k = (max_steps - launched + gs - 1) / gs;
        if (k <= 0) break;
        if (b->engine.tuning()) k = 1;     // one timed replay per height trial
        const int rc = b->engine.replay((int)k, why);
        if (rc) { drop_all(); return bfail(b, rc, why); }
        launched += k * gs;
        bool any = false;
        for (int i = 0; i < n && !any; ++i) any = !b->engine.finished(slots[i]);
        if (!any) break;
        per_check = std::min(per_check * 2, 16);
    }
    int status = SFV_OK;
    for (int i = 0; i < n; ++i) {
        const sfv::StepCtl& c = b->engine.ctl(slots[i]);
        t_io[i] = c.t;
        if (steps_out) steps_out[i] = c.steps;
        if (c.nonfinite && status == SFV_OK)
            status = bfail(b, SFV_ERR_NONFINITE_DT, "member " + std::to_string(i) + ": a time step produced a non-finite dt");
    }
    drop_all();
    return status;
}

}  // extern "C"
