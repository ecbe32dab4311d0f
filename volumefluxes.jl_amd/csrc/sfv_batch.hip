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
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sinfvm_hip.h"
#include "sfv_handle.hpp"

using namespace sfv;

struct sfv_batch {
    std::vector<sfv_handle*> members;
    int device = 0;
    int precision = SFV_F64;
    hipStream_t stream = nullptr;
    BatchMember* d_mem = nullptr;
    BatchHeader* d_hdr = nullptr;
    BatchMember* h_mem = nullptr;      // pinned staging
    BatchHeader* h_hdr = nullptr;
    StepCtl* d_ctl = nullptr;
    StepCtl* h_ctl = nullptr;          // pinned mirror
    int capacity = 0;
    int blocks_stage = 0, blocks_bc = 0;
    int graph_steps = 8;               // even: launch parity restarts at 0 on every replay
    hipGraphExec_t graph = nullptr;
    std::string err;

    ~sfv_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        if (d_mem) (void)hipFree(d_mem);
        if (d_hdr) (void)hipFree(d_hdr);
        if (d_ctl) (void)hipFree(d_ctl);
        if (h_mem) (void)hipHostFree(h_mem);
        if (h_hdr) (void)hipHostFree(h_hdr);
        if (h_ctl) (void)hipHostFree(h_ctl);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

thread_local std::string g_batch_error;

int bfail(sfv_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    else g_batch_error = msg;
    return code;
}

#define BHIP(b, expr)                                                                                \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return bfail((b), SFV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

// Capture graph_steps RK2 steps (stage 1, stage 2, boundary condition each),
// launch parity alternating 0, 1, ... on the batch stream.
int capture(sfv_batch* b) {
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int s = 0; s < b->graph_steps && e == hipSuccess; ++s) {
        const int lp = s & 1;
        if (b->precision == SFV_F64) {
            e = launch_march_batch<double>(b->stream, 1, b->d_mem, b->d_hdr, lp, b->blocks_stage);
            if (e == hipSuccess) e = launch_march_batch<double>(b->stream, 2, b->d_mem, b->d_hdr, lp, b->blocks_stage);
            if (e == hipSuccess) e = launch_bc_batch<double>(b->stream, b->d_mem, b->d_hdr, lp, b->blocks_bc);
        } else {
            e = launch_march_batch<float>(b->stream, 1, b->d_mem, b->d_hdr, lp, b->blocks_stage);
            if (e == hipSuccess) e = launch_march_batch<float>(b->stream, 2, b->d_mem, b->d_hdr, lp, b->blocks_stage);
            if (e == hipSuccess) e = launch_bc_batch<float>(b->stream, b->d_mem, b->d_hdr, lp, b->blocks_bc);
        }
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return bfail(b, SFV_ERR_HIP, std::string("batched launch during capture: ") + hipGetErrorString(e));
    }
    if (e2 != hipSuccess) return bfail(b, SFV_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
    const hipError_t e3 = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) {
        b->graph = nullptr;
        return bfail(b, SFV_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e3));
    }
    return SFV_OK;
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
        if (h->device != handles[0]->device)
            return bfail(nullptr, SFV_ERR_ARGUMENT, "all members must live on one device");
        if (h->cfg.precision != handles[0]->cfg.precision)
            return bfail(nullptr, SFV_ERR_ARGUMENT, "all members must have the same precision");
    }
    sfv_batch* b = new sfv_batch();
    b->members.assign(handles, handles + n);
    b->device = handles[0]->device;
    b->precision = handles[0]->cfg.precision;
    b->capacity = n;
    auto bail = [&](int code, const std::string& msg) {
        delete b;
        return bfail(nullptr, code, msg);
    };
    if (hipSetDevice(b->device) != hipSuccess) return bail(SFV_ERR_HIP, "hipSetDevice failed");
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(SFV_ERR_HIP, "hipStreamCreate failed");
    if (hipMalloc((void**)&b->d_mem, sizeof(BatchMember) * n) != hipSuccess ||
        hipMalloc((void**)&b->d_hdr, sizeof(BatchHeader)) != hipSuccess ||
        hipMalloc((void**)&b->d_ctl, sizeof(StepCtl) * n) != hipSuccess)
        return bail(SFV_ERR_OUT_OF_MEMORY, "hipMalloc of the batch tables failed");
    if (hipHostMalloc((void**)&b->h_mem, sizeof(BatchMember) * n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_hdr, sizeof(BatchHeader), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_ctl, sizeof(StepCtl) * n, hipHostMallocDefault) != hipSuccess)
        return bail(SFV_ERR_OUT_OF_MEMORY, "hipHostMalloc of the batch staging failed");
    std::memset(b->h_mem, 0, sizeof(BatchMember) * n);
    std::memset(b->h_ctl, 0, sizeof(StepCtl) * n);
    hipDeviceProp_t prop;
    int cus = 256;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
        cus = prop.multiProcessorCount;
    const int per_cu = std::min(march_batch_blocks_per_cu(b->precision, 1), march_batch_blocks_per_cu(b->precision, 2));
    b->blocks_stage = std::max(8, (cus * per_cu + 7) / 8 * 8);
    b->blocks_bc = std::max(1, cus);
    *out = b;
    return SFV_OK;
}

void sfv_batch_destroy(sfv_batch* b) { delete b; }

const char* sfv_batch_last_error(const sfv_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

int32_t sfv_batch_size(const sfv_batch* b) { return b ? (int32_t)b->members.size() : 0; }

int sfv_batch_run(sfv_batch* b, const double* t_end, int32_t match_endtime, double maximum_timestep,
                  int64_t max_steps, double* t_io, int64_t* steps_out) {
    if (!b) return SFV_ERR_ARGUMENT;
    if (!t_io) return bfail(b, SFV_ERR_ARGUMENT, "t_io is NULL");
    if (!t_end && max_steps <= 0) return bfail(b, SFV_ERR_ARGUMENT, "give t_end or a positive max_steps");
    const int n = (int)b->members.size();
    for (int i = 0; i < n; ++i)
        if (!b->members[i]->has_state) return bfail(b, SFV_ERR_STATE, "member " + std::to_string(i) + " has no state");
    BHIP(b, hipSetDevice(b->device));
    for (sfv_handle* h : b->members) BHIP(b, hipStreamSynchronize(h->stream));   // pending set_state etc.

    // member table in the batch order (tiles of member 0, then member 1, ...)
    int tiles = 0, lines = 0;
    for (int i = 0; i < n; ++i) {
        sfv_handle* h = b->members[i];
        BatchMember& M = b->h_mem[i];
        std::memset(&M, 0, sizeof(M));
        M.P = h->P;
        M.ctl = b->d_ctl + i;
        M.U[0] = h->U[0];
        M.U[1] = h->U[1];
        M.RHS = h->RHS;
        M.B = h->B;
        M.cur0 = (int)(h->cur & 1);
        M.slot0 = (int)(h->step_parity & 1);
        M.my = march_rows_per_wave(h->P, 1);
        const int segs = (h->P.ny + M.my - 1) / M.my;
        M.strips = (h->P.nx + MARCH_COLS - 1) / MARCH_COLS;
        M.seggroups = (segs + MARCH_WAVES - 1) / MARCH_WAVES;
        M.tile0 = tiles;
        M.bc0 = lines;
        tiles += M.strips * M.seggroups;
        lines += h->P.nx + h->P.ny;
        M.t0 = t_io[i];
        M.until_endtime = t_end ? 1 : 0;
        M.endtime = t_end ? t_end[i] : NAN;
        M.match_endtime = match_endtime ? 1 : 0;
        M.has_max_timestep = maximum_timestep > 0 ? 1 : 0;
        M.max_timestep = maximum_timestep > 0 ? maximum_timestep : 0.0;
        M.max_steps = max_steps;
        M.precision = b->precision;
    }
    b->h_hdr->members = n;
    b->h_hdr->tiles = tiles;
    b->h_hdr->bc_lines = lines;
    b->h_hdr->pad = 0;
    BHIP(b, hipMemcpyAsync(b->d_mem, b->h_mem, sizeof(BatchMember) * n, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d_hdr, b->h_hdr, sizeof(BatchHeader), hipMemcpyHostToDevice, b->stream));
    BHIP(b, launch_prepare_batch(b->stream, b->d_mem, b->d_hdr, n));
    if (!b->graph) {
        const int rc = capture(b);
        if (rc) return rc;
    }

    // Replay until every member has finished; a fixed step count needs no
    // look at the device on the way.
    int64_t launched = 0, replays_per_check = 1;
    for (;;) {
        int64_t k = replays_per_check;
        if (max_steps > 0) k = std::min<int64_t>(k, (max_steps - launched + b->graph_steps - 1) / b->graph_steps);
        if (!t_end) k = (max_steps - launched + b->graph_steps - 1) / b->graph_steps;
        if (k <= 0) break;
        for (int64_t r = 0; r < k; ++r) BHIP(b, hipGraphLaunch(b->graph, b->stream));
        launched += k * b->graph_steps;
        BHIP(b, hipMemcpyAsync(b->h_ctl, b->d_ctl, sizeof(StepCtl) * n, hipMemcpyDeviceToHost, b->stream));
        BHIP(b, hipStreamSynchronize(b->stream));
        bool any = false;
        for (int i = 0; i < n && !any; ++i) any = b->h_ctl[i].slot[(b->h_mem[i].slot0 + launched) & 1].active != 0;
        if (!any) break;
        replays_per_check = std::min<int64_t>(replays_per_check * 2, 16);
    }
    int status = SFV_OK;
    for (int i = 0; i < n; ++i) {
        sfv_handle* h = b->members[i];
        const StepCtl& c = b->h_ctl[i];
        h->cur = (h->cur + c.steps) & 1;
        h->step_parity = (h->step_parity + c.steps) & 1;
        if (c.steps > 0) { h->last_dt = c.dt; h->have_dt = true; }
        t_io[i] = c.t;
        if (steps_out) steps_out[i] = c.steps;
        if (c.nonfinite && status == SFV_OK) {
            status = bfail(b, SFV_ERR_NONFINITE_DT, "member " + std::to_string(i) + ": a time step produced a non-finite dt");
        }
    }
    return status;
}

}  // extern "C"
