// swe_kernels_ctl.hip — ghost-cell boundary condition and the device-side
// step control (time, dt, per-step flags) that lets a run advance without a
// host synchronisation per step.
#include "swe_kernels.hpp"

namespace sfv {


// update_bc! (bc.jl:29-110): X ghosts of every interior row, then Y ghosts of
// every interior column; corners are never written.  One thread per row or
// column reads all source cells first, then writes the ghosts, which gives
// the reference's sequential per-line semantics (KParams::src_* maps).
// Line tid < ny is interior row gc+tid, ny <= tid < ny+nx interior column
// gc+tid-ny.  With last_stage set, line 0 also closes the step: t += dt
// (simulator.jl:205), and arms the next step's slot.
template <class S>
__device__ __forceinline__ void bc_line(const KParams& P, StepCtl* ctl, int slot, S* __restrict__ U, int mode,
                                        int tid) {
    const int last_stage = mode == BC_LAST_STAGE;
    if (mode != BC_STANDALONE && !ctl->slot[slot].active) {
        // An inactive step passes inactivity on to the next slot, so steps
        // launched past the end of a run stay no-ops.
        if (last_stage && tid == 0) ctl->slot[1 - slot].active = 0;
        return;
    }
    const size_t n = (size_t)P.nxT * P.nyT;
    const int gc = P.gc;
    if (tid < P.ny) {
        const int j = gc + tid;
        S lo[MAX_GC][3], hi[MAX_GC][3];
        for (int g = 0; g < gc; ++g)
            for (int v = 0; v < 3; ++v) {
                lo[g][v] = U[v * n + cid(P, P.src_lo_x[g], j)];
                hi[g][v] = U[v * n + cid(P, P.src_hi_x[g], j)];
            }
        for (int g = 0; g < gc; ++g)
            for (int v = 0; v < 3; ++v) {
                const bool nl = P.flip_lo_x[g] && v == 1, nh = P.flip_hi_x[g] && v == 1;
                U[v * n + cid(P, g, j)] = nl ? (S)(-lo[g][v]) : lo[g][v];
                U[v * n + cid(P, P.nxT - gc + g, j)] = nh ? (S)(-hi[g][v]) : hi[g][v];
            }
    } else if (tid < P.ny + P.nx) {
        const int i = gc + (tid - P.ny);
        S lo[MAX_GC][3], hi[MAX_GC][3];
        for (int g = 0; g < gc; ++g)
            for (int v = 0; v < 3; ++v) {
                lo[g][v] = U[v * n + cid(P, i, P.src_lo_y[g])];
                hi[g][v] = U[v * n + cid(P, i, P.src_hi_y[g])];
            }
        for (int g = 0; g < gc; ++g)
            for (int v = 0; v < 3; ++v) {
                const bool nl = P.flip_lo_y[g] && v == 2, nh = P.flip_hi_y[g] && v == 2;
                U[v * n + cid(P, i, g)] = nl ? (S)(-lo[g][v]) : lo[g][v];
                U[v * n + cid(P, i, P.nyT - gc + g)] = nh ? (S)(-hi[g][v]) : hi[g][v];
            }
    }
    if (last_stage && tid == 0) close_step<S>(P, ctl, slot);
}

template <class S>
__global__ void __launch_bounds__(256) k_bc(KParams P, StepCtl* ctl, int slot, S* __restrict__ U, int mode) {
    bc_line<S>(P, ctl, slot, U, mode, blockIdx.x * blockDim.x + threadIdx.x);
}

// Initialise the step-control block for a host call.  until_endtime = 0 runs
// a fixed number of steps (endtime is then NaN and never ends the loop).
__device__ void prepare_ctl(const KParams& P, StepCtl* c, int slot, double t, double endtime, int match_endtime,
                            int has_max_timestep, double max_timestep, long long max_steps, int until_endtime,
                            int precision) {
    c->t = t;
    c->endtime = until_endtime ? endtime : __builtin_nan("");
    c->match_endtime = match_endtime;
    c->has_max_timestep = has_max_timestep;
    c->max_timestep = max_timestep;
    c->steps = 0;
    c->max_steps = max_steps;
    c->nonfinite = 0;
    c->precision = precision;
    bool act = true;
    if (until_endtime) act = t < endtime;
    for (int s = 0; s < 2; ++s) {
        c->slot[s].wsmax[0] = 0ull;
        c->slot[s].wsmax[1] = 0ull;
        c->slot[s].active = 0;
        set_rates(P, c->slot[s], t);
    }
    c->slot[slot].active = act ? 1 : 0;
}

__global__ void k_prepare(KParams P, StepCtl* ctl, int slot, double t, double endtime, int match_endtime,
                          int has_max_timestep, double max_timestep, long long max_steps, int until_endtime,
                          int precision) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    prepare_ctl(P, ctl, slot, t, endtime, match_endtime, has_max_timestep, max_timestep, max_steps, until_endtime,
                precision);
}

// Batched prepare: one thread per member, with the member's run parameters.
__global__ void k_prepare_batch(const BatchMember* __restrict__ mem, const BatchHeader* __restrict__ hdr) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= hdr->members) return;
    const BatchMember& M = mem[m];
    prepare_ctl(M.P, M.ctl, M.slot0 & 1, M.t0, M.endtime, M.match_endtime, M.has_max_timestep, M.max_timestep,
                M.max_steps, M.until_endtime, M.precision);
}

hipError_t launch_prepare_batch(hipStream_t st, const BatchMember* mem, const BatchHeader* hdr, int members) {
    hipLaunchKernelGGL(k_prepare_batch, dim3((members + 63) / 64), dim3(64), 0, st, mem, hdr);
    return hipPeekAtLastError();
}

template <class S>
hipError_t launch_bc(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, S* U, int mode) {
    const int nthreads = P.nx + P.ny;
    hipLaunchKernelGGL(k_bc<S>, dim3((nthreads + 255) / 256), dim3(256), 0, st, P, ctl, slot, U, mode);
    return hipPeekAtLastError();
}
hipError_t launch_prepare(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, double t, double endtime,
                          int match_endtime, int has_max_timestep, double max_timestep, long long max_steps,
                          int until_endtime, int precision) {
    hipLaunchKernelGGL(k_prepare, dim3(1), dim3(64), 0, st, P, ctl, slot, t, endtime, match_endtime, has_max_timestep,
                       max_timestep, max_steps, until_endtime, precision);
    return hipPeekAtLastError();
}
// A plain device copy of `words` 4-byte words (16 bytes per thread and
// iteration, grid-stride), used as the known-traffic reference kernel when
// HBM byte counters are calibrated.  Both pointers are hipMalloc results
// (256-byte aligned).
__global__ void __launch_bounds__(256) k_copy_words(const unsigned* __restrict__ src, unsigned* __restrict__ dst,
                                                    size_t words) {
    const size_t n16 = words / 4;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (size_t i = tid; i < n16; i += stride) d4[i] = s4[i];
    if (tid < words - 4 * n16) dst[4 * n16 + tid] = src[4 * n16 + tid];
}

hipError_t launch_copy(hipStream_t st, const void* src, void* dst, size_t bytes) {
    const size_t words = bytes / 4;
    const size_t n16 = words / 4;
    size_t blocks = (n16 + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_copy_words, dim3((unsigned)blocks), dim3(256), 0, st, (const unsigned*)src, (unsigned*)dst,
                       words);
    return hipPeekAtLastError();
}

template <class W>
__global__ void __launch_bounds__(256) k_interior_copy(int nx, int ny, int nxT, int gc, size_t ncell, W* packed,
                                                       W* state, int to_state) {
    const int i = blockIdx.x * 256 + threadIdx.x, v = blockIdx.z;
    if (i >= nx) return;
    for (int j = blockIdx.y; j < ny; j += gridDim.y) {
        const size_t a = (size_t)v * nx * ny + (size_t)j * nx + i;
        const size_t b = (size_t)v * ncell + (size_t)(j + gc) * nxT + (i + gc);
        if (to_state) state[b] = packed[a];
        else packed[a] = state[b];
    }
}

hipError_t launch_interior_copy(hipStream_t st, const KParams& P, const void* packed, void* state, size_t es,
                                int to_state) {
    if (P.nx <= 0 || P.ny <= 0) return hipSuccess;
    const dim3 grid((unsigned)((P.nx + 255) / 256), (unsigned)(P.ny < 16384 ? P.ny : 16384), 3);
    const size_t ncell = (size_t)P.nxT * P.nyT;
    if (es == 8)
        hipLaunchKernelGGL(k_interior_copy<unsigned long long>, grid, dim3(256), 0, st, P.nx, P.ny, P.nxT, P.gc,
                           ncell, (unsigned long long*)packed, (unsigned long long*)state, to_state);
    else if (es == 4)
        hipLaunchKernelGGL(k_interior_copy<unsigned>, grid, dim3(256), 0, st, P.nx, P.ny, P.nxT, P.gc, ncell,
                           (unsigned*)packed, (unsigned*)state, to_state);
    else
        return hipErrorInvalidValue;
    return hipPeekAtLastError();
}

template hipError_t launch_bc<double>(hipStream_t, const KParams&, StepCtl*, int, double*, int);
template hipError_t launch_bc<float>(hipStream_t, const KParams&, StepCtl*, int, float*, int);

}  // namespace sfv
