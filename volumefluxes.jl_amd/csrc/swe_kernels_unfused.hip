// swe_kernels_unfused.hip — the per-pass kernels of one time step, one thread
// per cell.  They keep the reference's pass boundaries (flux/sources → RK
// update + friction + clamp → boundary condition) and are the device-side
// parity reference for the fused stage kernels (swe_kernels_fused.hip); the
// Forward-Euler stepper and the boundary condition always use them.
#include "swe_kernels.hpp"

namespace sfv {

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Block max of two values, then one atomic per block into the step slot.
// Speeds are >= +0.0 or NaN, whose bit patterns order like the values (and a
// NaN orders above +Inf), so an unsigned max of the bits is max() with NaN
// propagation (numericalflux.jl:32).
__device__ void block_max_to_slot(unsigned long long sx, unsigned long long sy, StepSlot* slot) {
    __shared__ unsigned long long red[2][16];
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int lane = tid & 63, wid = tid >> 6;
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    sx = wave_max_u64(sx);
    sy = wave_max_u64(sy);
    if (lane == 0) { red[0][wid] = sx; red[1][wid] = sy; }
    __syncthreads();
    if (tid == 0) {
        unsigned long long mx = red[0][0], my = red[1][0];
        for (int w = 1; w < nw; ++w) { mx = red[0][w] > mx ? red[0][w] : mx; my = red[1][w] > my ? red[1][w] : my; }
        if (mx) atomicMax(&slot->wsmax[0], mx);
        if (my) atomicMax(&slot->wsmax[1], my);
    }
}

// ------------------------------------------------------------ flux (+sources)
// add_time_derivative! (system.jl:62-75) for every cell: interior cells get
// the X then Y {flux difference, bottom source} and the non-directional
// sources; ghost cells get only the non-directional sources (rain and
// infiltration run over all cells).  RHS has the state's storage type: the
// reference accumulates into a realtype volume and rounds at every store.
template <class S>
__global__ void __launch_bounds__(256) k_flux(KParams P, StepCtl* ctl, int slot, const S* __restrict__ X,
                                              S* __restrict__ RHS, const double* __restrict__ B, int record) {
    if (!ctl->slot[slot].active) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // the launchers' blocks are 64 wide, so a wave shares threadIdx.y
    const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * blockDim.y + threadIdx.y);
    const size_t n = (size_t)P.nxT * P.nyT;
    unsigned long long sx = 0, sy = 0;
    if (i < P.nxT && j < P.nyT) {
        const size_t k = cid(P, i, j);
        S acc[3] = {S(0), S(0), S(0)};
        const bool interior = i >= P.gc && i < P.nxT - P.gc && j >= P.gc && j < P.nyT - P.gc;
        if (interior) {
#pragma unroll 1
            for (int d = 0; d < 2; ++d) {
                const size_t st = d == 0 ? 1 : (size_t)P.nxT;
                const int di = d == 0 ? 1 : 0, dj = d == 0 ? 0 : 1;
                // cells k-2..k+2 along d; without reconstruction only k-1..k+1
                // are read (gc may then be 1)
                S c[5][3];
                const size_t kd = (size_t)(k - 2 * st);
#pragma unroll
                for (int o = 0; o < 5; ++o) {
                    const bool need = P.rec == 1 || (o >= 1 && o <= 3);
#pragma unroll
                    for (int v = 0; v < 3; ++v) c[o][v] = need ? X[v * n + kd + (size_t)o * st] : S(0);
                }
                // face states of cells k-1, k, k+1 in this direction
                S Lf[3][3], Rf[3][3];
#pragma unroll
                for (int o = 0; o < 3; ++o) {
                    const int ii = i + (o - 1) * di, jj = j + (o - 1) * dj;
                    if (P.rec == 1) {
                        reconstruct_cell(c[o], c[o + 1], c[o + 2], P.theta, B_face_left(P, B, ii, jj, d),
                                         B_face_right(P, B, ii, jj, d), Lf[o], Rf[o]);
                    } else {
                        const double bc = B_cell(P, B, ii, jj);
#pragma unroll
                        for (int v = 0; v < 3; ++v) { Lf[o][v] = c[o + 1][v]; Rf[o][v] = c[o + 1][v]; }
                        Lf[o][0] = (S)(c[o + 1][0] - bc);
                        Rf[o][0] = (S)(c[o + 1][0] - bc);
                    }
                }
                double Fp[3], Fm[3];
                const double sp = central_upwind(d, Rf[1], Lf[2], P.phys, Fp);
                const double sm = central_upwind(d, Rf[0], Lf[1], P.phys, Fm);
                const double inv = d == 0 ? P.inv_dx : P.inv_dy;
#pragma unroll
                for (int v = 0; v < 3; ++v) acc[v] = (S)(acc[v] - inv * (Fp[v] - Fm[v]));
                const S ws = (S)jl_max(sp, sm);
                const unsigned long long bits = (unsigned long long)__double_as_longlong((double)ws);
                if (d == 0) sx = bits; else sy = bits;
                if (P.bottom) {
                    const double Br = B_face_right(P, B, i, j, d), Bl = B_face_left(P, B, i, j, d);
                    const double dd = d == 0 ? P.dx : P.dy;
                    const S hsum = Rf[1][0] + Lf[1][0];
                    acc[1 + d] = (S)(acc[1 + d] + -P.phys.g * slope_div(Br - Bl, dd, inv, d == 0 ? P.dx_pow2 : P.dy_pow2) * (hsum / 2.0));
                }
            }
        }
        acc[0] = apply_sources(P, ctl->slot[slot], acc[0], k);
#pragma unroll
        for (int v = 0; v < 3; ++v) RHS[v * n + k] = acc[v];
    }
    if (record) block_max_to_slot(sx, sy, &ctl->slot[slot]);
}

// -------------------------------------------------- RK update + friction + clamp
// rungekutta.jl:38-49 / forwardeuler.jl:41-43, then implicit_substep!
// (friction.jl:80-84, from the substep input X) and post_proc_substep!
// (timestepper.jl:53-63), over all cells.  stage 1 computes dt
// (simulator.jl:147-151); stage 2 averages with U0.  With rhs_interior set,
// RHS holds only the interior cells (the march stage-1 kernel's output) and
// the other cells take their right-hand side as k_flux gives it there: the
// non-directional sources on w, zero on hu and hv; with ghost_rhs given,
// that right-hand side is also stored there (ghost_rhs may be RHS itself:
// RHS is read only at interior cells, ghost_rhs written only at the others).
template <class S>
__global__ void __launch_bounds__(256) k_finish(KParams P, StepCtl* ctl, int slot, int stage, const S* __restrict__ U0,
                                                const S* __restrict__ X, const S* RHS, S* __restrict__ Out,
                                                const double* __restrict__ B, int rhs_interior, int close,
                                                S* ghost_rhs) {
    const StepSlot& sl = ctl->slot[slot];
    const bool closer = close && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0;
    if (!sl.active) {
        if (closer) ctl->slot[1 - slot].active = 0;   // steps past the end of a run stay no-ops
        return;
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // the launchers' blocks are 64 wide, so a wave shares threadIdx.y
    const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * blockDim.y + threadIdx.y);
    double dt_d;
    S dt_s;
    if (stage == 1) {
        dt_d = stage1_dt<S>(P, ctl, sl);
        dt_s = (S)dt_d;
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
            ctl->dt = (double)dt_s;
            ctl->dt_stage1 = dt_d;
            // a one-stage step may close here (t += dt, next slot armed): the
            // other threads read only this step's slot
            if (close) close_step<S>(P, ctl, slot);
        }
    } else {
        dt_s = (S)ctl->dt;
        dt_d = (double)dt_s;
    }
    if (i >= P.nxT || j >= P.nyT) return;
    const size_t n = (size_t)P.nxT * P.nyT;
    const size_t k = cid(P, i, j);
    S x[3], o[3], r[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) x[v] = X[v * n + k];
    const bool interior = i >= P.gc && i < P.nxT - P.gc && j >= P.gc && j < P.nyT - P.gc;
    if (rhs_interior && !interior) {
        r[0] = apply_sources(P, ctl->slot[slot], S(0), k);
        r[1] = S(0);
        r[2] = S(0);
        if (ghost_rhs)
#pragma unroll
            for (int v = 0; v < 3; ++v) ghost_rhs[v * n + k] = r[v];
    } else {
#pragma unroll
        for (int v = 0; v < 3; ++v) r[v] = RHS[v * n + k];
    }
    if (stage == 1) {
#pragma unroll
        for (int v = 0; v < 3; ++v) o[v] = (S)(x[v] + dt_d * r[v]);
    } else {
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            o[v] = (S)(x[v] + dt_s * r[v]);
            o[v] = (S)(0.5 * (U0[v * n + k] + o[v]));
        }
    }
    const double Bm = B_cell(P, B, i, j);
    if (P.friction) {
        const double fs = friction_scalar(P.friction, P.Cz, P.phys.kappa, x[0], x[1], x[2], Bm);
        apply_friction(o, dt_s, fs);
    }
    if (o[0] - Bm < P.phys.cutoff) {
        o[0] = (S)jl_max((double)o[0], Bm);
        o[1] = (S)0.0;
        o[2] = (S)0.0;
    }
#pragma unroll
    for (int v = 0; v < 3; ++v) Out[v * n + k] = o[v];
}

template <class S>
hipError_t launch_flux(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, const S* X, S* RHS, const double* B,
                       int record) {
    const dim3 block(64, 4);
    const dim3 grid((P.nxT + 63) / 64, (P.nyT + 3) / 4);
    hipLaunchKernelGGL(k_flux<S>, grid, block, 0, st, P, ctl, slot, X, RHS, B, record);
    return hipPeekAtLastError();
}
template <class S>
hipError_t launch_finish(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, int stage, const S* U0, const S* X,
                         const S* RHS, S* Out, const double* B, int rhs_interior, int close, S* ghost_rhs) {
    const dim3 block(64, 4);
    const dim3 grid((P.nxT + 63) / 64, (P.nyT + 3) / 4);
    hipLaunchKernelGGL(k_finish<S>, grid, block, 0, st, P, ctl, slot, stage, U0, X, RHS, Out, B, rhs_interior,
                       close, ghost_rhs);
    return hipPeekAtLastError();
}
template hipError_t launch_flux<double>(hipStream_t, const KParams&, StepCtl*, int, const double*, double*,
                                        const double*, int);
template hipError_t launch_flux<float>(hipStream_t, const KParams&, StepCtl*, int, const float*, float*,
                                       const double*, int);
template hipError_t launch_finish<double>(hipStream_t, const KParams&, StepCtl*, int, int, const double*,
                                          const double*, const double*, double*, const double*, int, int, double*);
template hipError_t launch_finish<float>(hipStream_t, const KParams&, StepCtl*, int, int, const float*, const float*,
                                         const float*, float*, const double*, int, int, float*);

template __global__ void k_flux<double>(KParams, StepCtl*, int, const double*, double*, const double*, int);
template __global__ void k_flux<float>(KParams, StepCtl*, int, const float*, float*, const double*, int);
template __global__ void k_finish<double>(KParams, StepCtl*, int, int, const double*, const double*, const double*,
                                          double*, const double*, int, int, double*);
template __global__ void k_finish<float>(KParams, StepCtl*, int, int, const float*, const float*, const float*,
                                         float*, const double*, int, int, float*);

}  // namespace sfv
