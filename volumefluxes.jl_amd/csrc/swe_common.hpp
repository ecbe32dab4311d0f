// swe_common.hpp — kernel parameter block, device step-control block and
// indexing shared by the gfx950 kernels and the host side of the library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swe_physics.hpp"

namespace sfv {

constexpr int MAX_GC = 8;

// Everything a kernel needs that is fixed for the life of a handle.
struct KParams {
    int nx, ny, gc, nxT, nyT;
    double dx, dy, inv_dx, inv_dy;  // inv_d is 1/Δ as numericalflux.jl:27 computes it
    int dx_pow2, dy_pow2;           // Δ is a power of two (x/Δ == x*(1/Δ) exactly)
    PhysParams phys;
    double theta, cfl, Cz;
    int rec;        // 0 none, 1 linear
    int friction;   // 0 none, 1 bh2021, 2 bsa2012, 3 fcg2016
    int bottom;     // SourceTermBottom present
    int rain_kind;  // 0 none, 1 constant, 2 table, 3 per-cell field (FunctionalRain)
    int infil_kind; // 0 none, 1 constant, 2 horton
    int rain_first;
    int stepper;    // 1 FE, 2 RK2
    int rows;       // rows per wave segment of the march kernels (0: march_rows_per_wave's choice)
    int rain_n;
    const double* rain_rates;  // device (reference unit)
    const double* rain_times;  // device
    double rain_const;         // ConstantRain rate / 3600
    double infil_rate, f0, fc, hk;
    const double* factor;      // device (nxT*nyT) Horton factor field
    const double* rain_field;  // device (nxT*nyT) FunctionalRain rates of the current step
    int bcell_1d;   // B_cell = 0.5*(B[i,j] + B[i+1,j]) (a CartesianGrid{1} run embedded in (nx, 3))
    int ranged;     // θ, ρ, g, depth_cutoff and κ allow the ranged face evaluation (central_upwind_rng)
    int bc;
    // Ghost-cell source maps for the boundary condition, by side: low-side
    // ghost g is cell g, high-side ghost g is cell N-gc+g (N = nxT or nyT);
    // it takes the pre-BC value of cell src (same row/column), with the
    // normal momentum negated if flip is set.  Built on the host by
    // replaying bc.jl's per-ghost assignments in order, so order-dependent
    // maps (NeumannBC's high side, grids narrower than gc) are exact.
    int src_lo_x[MAX_GC], src_hi_x[MAX_GC], src_lo_y[MAX_GC], src_hi_y[MAX_GC];
    int flip_lo_x[MAX_GC], flip_hi_x[MAX_GC], flip_lo_y[MAX_GC], flip_hi_y[MAX_GC];
};

// Per-step flags live in two slots indexed by step parity, so the kernel that
// finishes step n can arm step n+1 while no kernel of step n reads that slot.
// One step's control: the kernels of a step read only their own slot (and
// the run-wide fields of StepCtl), and the step's last kernel arms the other
// slot for the next step, so the step can close while its kernels still run.
struct StepSlot {
    int32_t active;
    int32_t pad;
    unsigned long long wsmax[2];   // bit patterns of the non-negative max speeds (x, y)
    double t;             // the step's start time (realtype value)
    double rain_now;      // compute_rain(rain, t)
    double infil_tfac;    // fc + (f0 - fc)*exp(-k t) for Horton, the rate for constant
};

struct StepCtl {
    StepSlot slot[2];
    double t;             // current time (realtype value)
    double dt;            // current_timestep (realtype value)
    double dt_stage1;     // Float64 dt of stage 1 (== dt for Float64)
    double endtime;
    double max_timestep;
    int32_t match_endtime;
    int32_t has_max_timestep;
    int64_t steps;        // steps completed in this call
    int64_t max_steps;    // <= 0: unlimited
    int32_t nonfinite;    // a step produced a NaN/Inf dt
    int32_t precision;
};

// One member of a batch as the batched kernels see it (a device array of
// these, one per member, in the batch's tile order).  A launch with parity lp
// works on state buffer U[cur0 ^ lp] and step slot slot0 ^ lp of the member,
// so a fixed sequence of launches (a captured graph with an even number of
// steps) serves any membership: the host rewrites the table between replays.
struct BatchMember {
    KParams P;
    StepCtl* ctl;
    void* U[2];           // the state ping-pong pair
    void* RHS;            // stage-1 right-hand side
    const double* B;
    int cur0, slot0;
    int my;               // rows per wave segment
    int strips, segs;     // wave strips and segments per strip
    int tile0;            // first block tile of this member in the batch order
    int ntiles;           // its block tiles: march_blocks(strips, segs)
    int pad;
    // run parameters, read by the batched prepare
    double t0, endtime, max_timestep;
    long long max_steps;
    int match_endtime, has_max_timestep, until_endtime, precision;
};
// Batch-wide counts, in device memory beside the member table.
struct BatchHeader {
    int members;
    int tiles;            // sum of the members' ntiles
    int prio;             // 1: the waves take the highest issue priority (a GUEST engine); 2: their
                          // rows-left priority stays below it (a HOST engine beside a guest); 0: alone
    int pad;
};

// The member owning tile t of the batch order: the last member whose first
// tile (tile0) is <= t.  Every member has at least one tile, so the first
// tiles increase strictly.
__device__ __forceinline__ int batch_find(const BatchMember* __restrict__ mem, int members, int t) {
    int lo = 0, hi = members - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int first = mem[mid].tile0;
        if (first <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ size_t cid(const KParams& P, int i, int j) { return (size_t)i + (size_t)P.nxT * j; }

__device__ __forceinline__ double Bcorner(const KParams& P, const double* B, int i, int j) {
    return B[(size_t)i + (size_t)(P.nxT + 1) * j];
}
// B_cell from the four corners a = B[i,j], b = B[i+1,j], c = B[i,j+1],
// d = B[i+1,j+1] in the reference's summation order (bottom_topography.jl:97),
// or the 1D lookup 0.5*(B[i] + B[i+1]) (bottom_topography.jl:92) when the
// grid carries a 1D run (bottom replicated along y).
__device__ __forceinline__ double bcell4(const KParams& P, double a, double b, double c, double d) {
    return P.bcell_1d ? 0.5 * (a + b) : 0.25 * (((a + b) + c) + d);
}
__device__ __forceinline__ double B_cell(const KParams& P, const double* B, int i, int j) {
    return bcell4(P, Bcorner(P, B, i, j), Bcorner(P, B, i + 1, j), Bcorner(P, B, i, j + 1), Bcorner(P, B, i + 1, j + 1));
}
__device__ __forceinline__ double B_face_left(const KParams& P, const double* B, int i, int j, int d) {
    return d == 0 ? 0.5 * (Bcorner(P, B, i, j) + Bcorner(P, B, i, j + 1))
                  : 0.5 * (Bcorner(P, B, i, j) + Bcorner(P, B, i + 1, j));
}
__device__ __forceinline__ double B_face_right(const KParams& P, const double* B, int i, int j, int d) {
    return d == 0 ? 0.5 * (Bcorner(P, B, i + 1, j) + Bcorner(P, B, i + 1, j + 1))
                  : 0.5 * (Bcorner(P, B, i, j + 1) + Bcorner(P, B, i + 1, j + 1));
}

// source_term_rain.jl:45-54 scan of a TimeDependentRain table.
__device__ inline double compute_rain(const KParams& P, double t) {
    if (P.rain_kind == 1) return P.rain_const;
    if (P.rain_kind == 2) {
        int index = 0;
        while (index + 1 < P.rain_n) {
            if (t < P.rain_times[index + 1]) break;
            index += 1;
        }
        return P.rain_rates[index] / 3600;
    }
    return 0.0;
}

// Non-directional sources on the w component, in source-list order
// (system.jl:72-74): rain adds rate(t), infiltration subtracts its rate.
template <class S>
__device__ __forceinline__ S apply_sources(const KParams& P, const StepSlot& sl, S w, size_t k) {
    auto infil = [&](S x) -> S {
        if (P.infil_kind == 2) return (S)(x - P.factor[k] * sl.infil_tfac);
        if (P.infil_kind == 1) return (S)(x - sl.infil_tfac);
        return x;
    };
    auto rain = [&](S x) -> S {
        if (P.rain_kind == 3) return (S)(x + P.rain_field[k]);
        return P.rain_kind ? (S)(x + sl.rain_now) : x;
    };
    return P.rain_first ? infil(rain(w)) : rain(infil(w));
}

// The time-dependent source rates of a step starting at t.
__device__ inline void set_rates(const KParams& P, StepSlot& s, double t) {
    s.t = t;
    s.rain_now = compute_rain(P, t);
    if (P.infil_kind == 2) s.infil_tfac = P.fc + (P.f0 - P.fc) * exp(-P.hk * t);
    else if (P.infil_kind == 1) s.infil_tfac = P.infil_rate;
    else s.infil_tfac = 0.0;
}

// Close the step of `slot` with its dt: t += dt (simulator.jl:205), count
// it, and arm the next slot (active unless the run ends here).  One thread.
template <class S, bool NO_INFIL = false>
__device__ inline void close_step(const KParams& P, StepCtl* c, int slot) {
    const S dt = (S)c->dt;
    const S t = (S)((S)c->slot[slot].t + dt);
    c->t = (double)t;
    c->steps += 1;
    if (!__builtin_isfinite((double)dt)) c->nonfinite = 1;
    bool act = c->max_steps <= 0 || c->steps < c->max_steps;
    if (c->endtime == c->endtime) act = act && (t < c->endtime);   // NaN endtime: fixed step count
    StepSlot& nx = c->slot[1 - slot];
    nx.wsmax[0] = 0ull;
    nx.wsmax[1] = 0ull;
    nx.active = act ? 1 : 0;
    if (NO_INFIL) {             // the caller knows the run has no infiltration
        nx.t = (double)t;
        nx.rain_now = compute_rain(P, (double)t);
        nx.infil_tfac = 0.0;
    } else {
        set_rates(P, nx, (double)t);
    }
}


// The same, out of line, for the persistent batch kernels (their KParams
// live in the member table): inlined into the tile loop, the rate code
// costs the whole loop registers.
template <class S, bool NO_INFIL = false>
__device__ __attribute__((noinline)) void close_step_call(const KParams& P, StepCtl* c, int slot) {
    close_step<S, NO_INFIL>(P, c, slot);
}

// timestep_computer (simulator.jl:147-150) with the max_dt rule of
// simulate_to_time (simulator.jl:200-203).
template <class S>
__device__ __forceinline__ double stage1_dt(const KParams& P, const StepCtl* ctl, const StepSlot& sl) {
    double max_dt = ctl->match_endtime ? ctl->endtime - (S)sl.t : __builtin_inf();
    if (ctl->has_max_timestep) max_dt = jl_min(max_dt, ctl->max_timestep);
    const double ax = __longlong_as_double((long long)sl.wsmax[0]);
    const double ay = __longlong_as_double((long long)sl.wsmax[1]);
    const S cfl = (S)P.cfl;
    return jl_min(cfl * jl_min(P.dx / (S)ax, P.dy / (S)ay), max_dt);
}

}  // namespace sfv
