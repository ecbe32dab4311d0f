// swe_kernels.hpp — declarations of the gfx950 kernels of one time step.
#pragma once

#include "swe_common.hpp"

namespace sfv {

// ---- unfused per-pass kernels (swe_kernels_unfused.hip)
template <class S>
__global__ void k_flux(KParams P, StepCtl* ctl, int slot, const S* X, S* RHS, const double* B, int record);
template <class S>
__global__ void k_finish(KParams P, StepCtl* ctl, int slot, int stage, const S* U0, const S* X, const S* RHS, S* Out,
                         const double* B, int rhs_interior, int close, S* ghost_rhs);

// ---- boundary condition and step control (swe_kernels_ctl.hip)
template <class S>
__global__ void k_bc(KParams P, StepCtl* ctl, int slot, S* U, int mode);
__global__ void k_prepare(KParams P, StepCtl* ctl, int slot, double t, double endtime, int match_endtime,
                          int has_max_timestep, double max_timestep, long long max_steps, int until_endtime,
                          int precision);

// ---- fused stage kernels (swe_kernels_fused.hip)
// Tile of interior cells per block and threads per block.
constexpr int FT_X = 64;
constexpr int FT_Y = 8;
constexpr int F_THREADS = 256;
template <class S>
__global__ void k_stage1(KParams P, StepCtl* ctl, int slot, const S* U0, S* RHS, const double* B);
template <class S>
__global__ void k_stage2(KParams P, StepCtl* ctl, int slot, const S* U0, const S* RHS1, S* U2, const double* B);

// ---- host launchers (one per kernel, defined beside it); each returns the
// launch status.  BC modes: 0 inside a step, 1 last stage of a step (closes
// the step), 2 standalone (set_current_state!, no step-control check).
enum { BC_IN_STEP = 0, BC_LAST_STAGE = 1, BC_STANDALONE = 2 };
template <class S>
hipError_t launch_flux(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, const S* X, S* RHS, const double* B,
                       int record);
template <class S>
hipError_t launch_finish(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, int stage, const S* U0, const S* X,
                         const S* RHS, S* Out, const double* B,
                         int rhs_interior = 0, int close = 0, S* ghost_rhs = nullptr);
template <class S>
hipError_t launch_bc(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, S* U, int mode);
hipError_t launch_prepare(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, double t, double endtime,
                          int match_endtime, int has_max_timestep, double max_timestep, long long max_steps,
                          int until_endtime, int precision);
// Batched kernels over a device member table (BatchMember[hdr->members]).
// Stage kernels are persistent: `blocks` workgroups walk the batch's tiles.
// fast: every member satisfies march_fast (the specialised kernels).
template <class S>
hipError_t launch_march_batch(hipStream_t st, int stage, const BatchMember* mem, const BatchHeader* hdr, int lp,
                              int blocks, bool fast);
// The configurations the specialised march kernels cover.
bool march_fast(const KParams& P);
hipError_t launch_prepare_batch(hipStream_t st, const BatchMember* mem, const BatchHeader* hdr, int members);
// Resident workgroups per CU of the batched stage kernel (occupancy).
int march_batch_blocks_per_cu(int precision, int stage);
int march_rows_per_wave(const KParams& P, int stage, int precision);   // precision: 0 f64, 1 f32
constexpr int MARCH_COLS = 60;   // output columns per wave strip
#ifndef SFV_MARCH_WAVES
#define SFV_MARCH_WAVES 4
#endif
constexpr int MARCH_WAVES = SFV_MARCH_WAVES;   // waves (segments) per workgroup
#ifndef SFV_MARCH_LINEAR
#define SFV_MARCH_LINEAR 1
#endif
// Workgroups of a march launch (or of one batch member) over strips x segs
// wave segments.  With SFV_MARCH_LINEAR the waves are numbered across the
// launch (strips fastest), so only the last workgroup has idle waves;
// otherwise a workgroup is MARCH_WAVES consecutive segments of one strip,
// and each strip's segments are padded to a multiple of MARCH_WAVES.
__host__ __device__ inline int march_blocks(int strips, int segs) {
    return SFV_MARCH_LINEAR ? (strips * segs + MARCH_WAVES - 1) / MARCH_WAVES
                            : strips * ((segs + MARCH_WAVES - 1) / MARCH_WAVES);
}

// Plain copy of `bytes` (a multiple of 4) from src to dst.
hipError_t launch_copy(hipStream_t st, const void* src, void* dst, size_t bytes);
// Interior cells between a contiguous (nx, ny, 3) array and the padded
// (nxT, nyT, 3) state: to_state = 1 scatters packed -> state, 0 gathers
// state -> packed.  es = element size (4 or 8).
hipError_t launch_interior_copy(hipStream_t st, const KParams& P, const void* packed, void* state, size_t es,
                                int to_state);
template <class S>
hipError_t launch_stage1(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, const S* U0, S* RHS,
                         const double* B);
// stage = 1: U0 -> OUT = RHS1 (+ speeds); stage = 2: U0, RHS1 -> OUT = U2
// (interior and domain corners).  Stage 2 reads halo ghost cells through the
// boundary-condition maps; stage 1 too with remap_halo = 1 (only when the
// maps name interior cells), so U2's ghosts are left unwritten and filled
// when something outside the march path needs them.  close = 1: stage 2 also
// closes the step (t += dt, next slot armed), as k_bc's last stage does.
template <class S>
hipError_t launch_march(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, int stage, const S* U0,
                        const S* RHS1, S* OUT, const double* B, int close = 0, int remap_halo = 0);
template <class S>
hipError_t launch_stage2(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, const S* U0, const S* RHS1, S* U2,
                         const double* B, int close = 0);

}  // namespace sfv
