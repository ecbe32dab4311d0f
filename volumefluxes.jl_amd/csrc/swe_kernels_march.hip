// swe_kernels_march.hip — RK2 stage kernels, one wavefront per strip.
//
// A wave owns MX = 60 output columns (lanes 2..61 of its 64; lanes 0, 1,
// 62, 63 are the x halo) and a segment of MY rows, and marches up the
// segment one row per iteration.  x neighbours come from the adjacent lanes
// (cross-lane shuffles), y neighbours from a sliding window of rows held in
// registers, so every cell is loaded once per wave, every reconstruction
// (linearreconstruction.jl:58-102) is evaluated once per cell and direction,
// and every face flux (centralupwind.jl:48-77) once per face — the reference
// evaluates each face from both sides with identical arguments
// (numericalflux.jl:25-26), so sharing it is bit-exact.  No LDS and no
// block-level barriers: waves are independent.
//
// The segment height is chosen per grid by the host so that the waves fill
// the GPU's resident wave slots in whole rounds.  Per iteration r the wave
// takes row jr = j0-2+r (its loads were issued one iteration earlier) and
//   1. y-reconstructs row jr-1 from rows jr-2..jr,
//   2. evaluates the y face between rows jr-2 and jr-1,
//   3. finishes row jr-2 (its x part was done one iteration earlier, its
//      y faces are now both known): flux difference in the reference's
//      X-then-Y order, bottom source, rain/infiltration, then either stores
//      the stage right-hand side (stage 1) or the RK2 result (stage 2),
//   4. x-reconstructs row jr-1 and evaluates its x faces (the x part).
// Stage 2 rebuilds U1 = clamp(friction(U0 + dt*RHS1)) for every loaded cell
// (ghost cells through the boundary-condition source maps), so the stage-1
// state never touches HBM.
#include <cmath>
#include <mutex>
#include <type_traits>
#include <utility>

#include "swe_kernels.hpp"

namespace sfv {

// Build-time tuning knobs (defaults are the measured best on MI355X):
//   SFV_MARCH_PREFETCH    issue the next row's loads one iteration early
//   SFV_MARCH_BRANCHFREE  evaluate the x and y faces with the branch-free forms
//   SFV_MARCH_WETFAST     straight-line face evaluation for all-wet waves
//   SFV_MARCH_ROWS        rows per wave segment (0: chosen per grid)
//   SFV_MARCH_MIN_WAVES   minimum waves per SIMD the register allocation must allow
//                         (SFV_MARCH_MIN_WAVES1: the same for stage 1 only)
//   SFV_MARCH_XCD         block order: 1 = each XCD gets a contiguous run of tiles,
//                         segments fastest (vertical halos stay in one L2);
//                         2 = same, strips fastest; 0 = hardware order
//   SFV_MARCH_UNROLL      unroll factor of the row loop (SFV_MARCH_UNROLL1 / 2: per stage)
//   SFV_MARCH_STARTFETCH  rows (0..4) whose loads a wave issues together before its first row
//   SFV_MARCH_KEEPU0      stage 2 keeps U0 of the last two rows in registers instead of reloading it
//   SFV_MARCH_FAST        specialised kernels for the common configuration (march_fast)
//   SFV_MARCH_PRIO        1 = waves with more rows left issue first; 0 = hardware order
//                         (oldest wave first)
//   SFV_MARCH_PRIO_DIV    the priority steps down every 1/DIV of a wave's rows, over
//                         its last 3/DIV
//   SFV_MARCH_LINEAR      1 = waves numbered across the launch (no per-strip padding of
//                         segments to whole workgroups); 0 = a workgroup per 4 segments of a strip
//   SFV_MARCH_TIMELINE    measurement build (`make timeline`): the single-grid kernels record
//                         each wave's start, end, hardware slot and row start times
#ifndef SFV_MARCH_PREFETCH
#define SFV_MARCH_PREFETCH 0
#endif
#ifndef SFV_MARCH_PREFETCH1
#define SFV_MARCH_PREFETCH1 SFV_MARCH_PREFETCH
#endif
#ifndef SFV_MARCH_PREFETCH2
#define SFV_MARCH_PREFETCH2 SFV_MARCH_PREFETCH
#endif
#ifndef SFV_MARCH_BRANCHFREE
#define SFV_MARCH_BRANCHFREE 0
#endif
#ifndef SFV_MARCH_WETFAST
#define SFV_MARCH_WETFAST 0
#endif
#ifndef SFV_MARCH_ROWS
#define SFV_MARCH_ROWS 0
#endif
#ifndef SFV_MARCH_MIN_WAVES
#define SFV_MARCH_MIN_WAVES 3
#endif
#ifndef SFV_MARCH_MIN_WAVES1
#define SFV_MARCH_MIN_WAVES1 SFV_MARCH_MIN_WAVES
#endif
#ifndef SFV_MARCH_XCD
#define SFV_MARCH_XCD 2
#endif
#ifndef SFV_MARCH_FAST
#define SFV_MARCH_FAST 1
#endif
#ifndef SFV_MARCH_KEEPU0
#define SFV_MARCH_KEEPU0 1
#endif
#ifndef SFV_MARCH_UNROLL
#define SFV_MARCH_UNROLL 1
#endif
#ifndef SFV_MARCH_STARTFETCH
#define SFV_MARCH_STARTFETCH 3
#endif
#ifndef SFV_MARCH_UNROLL1
#define SFV_MARCH_UNROLL1 SFV_MARCH_UNROLL
#endif
#ifndef SFV_MARCH_PRIO
#define SFV_MARCH_PRIO 1
#endif
#ifndef SFV_MARCH_PRIO_DIV
#define SFV_MARCH_PRIO_DIV 8
#endif
#ifndef SFV_MARCH_UNROLL2
#define SFV_MARCH_UNROLL2 SFV_MARCH_UNROLL
#endif


#ifndef SFV_MARCH_TIMELINE
#define SFV_MARCH_TIMELINE 0
#endif
#if SFV_MARCH_TIMELINE
// Wave timeline of the last single-grid launch of each stage (the 100 MHz
// real-time counter): [stage][wave] = {start, end, HW_ID, XCC_ID << 32 |
// strip << 16 | segment}, and the start of each of the first 32 row
// iterations.  Read with sfv_probe_timeline (this build only).
constexpr int TL_WAVES = 16384, TL_ROWS = 32;
__device__ unsigned long long tl_wave[2][TL_WAVES][4];
__device__ unsigned long long tl_rows[2][TL_WAVES][TL_ROWS];
#endif

// s_setprio takes an immediate: one branch per level
__device__ __forceinline__ void set_issue_priority(int lvl) {
    switch (lvl) {
        case 0: __builtin_amdgcn_s_setprio(0); break;
        case 1: __builtin_amdgcn_s_setprio(1); break;
        case 2: __builtin_amdgcn_s_setprio(2); break;
        default: __builtin_amdgcn_s_setprio(3); break;
    }
}

template <class F, size_t... I>
__device__ __forceinline__ void unrolled_rows(F& body, int r, std::index_sequence<I...>) {
    (body(r + (int)I, std::true_type{}, std::true_type{}), ...);
}

constexpr int MW = 64;                // lanes
constexpr int MX = MARCH_COLS;        // output columns per wave
constexpr int MWAVES = MARCH_WAVES;   // waves per block

// Lane shifts by one across the whole wave with DPP (wave_shr:1 / wave_shl:1,
// VALU latency instead of an LDS round trip): lane i receives lane i-1
// (shfl_up1) or lane i+1 (shfl_dn1).  Lane 0 (shfl_up1) / lane 63
// (shfl_dn1) have no source lane and read zero; both are halo lanes whose
// shifted values feed no output (the one exception, the corner column lane
// 63 needs, is supplied separately by next_col).  With no old value to keep,
// each shift is a single DPP move into a fresh register.
#ifndef SFV_MARCH_DPP_ZERO
#define SFV_MARCH_DPP_ZERO 1
#endif
template <int CTRL>
__device__ __forceinline__ int dpp_shift(int v) {
#if SFV_MARCH_DPP_ZERO
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
#else
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xf, 0xf, false);
#endif
}
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_shift<CTRL>((int)(unsigned long long)b);
    const int hi = dpp_shift<CTRL>((int)((unsigned long long)b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float v) { return __int_as_float(dpp_shift<CTRL>(__float_as_int(v))); }
constexpr int DPP_WAVE_SHR1 = 0x138, DPP_WAVE_SHL1 = 0x130;
template <class T>
__device__ __forceinline__ T shfl_up1(T v) { return dpp_shift<DPP_WAVE_SHR1>(v); }
template <class T>
__device__ __forceinline__ T shfl_dn1(T v) { return dpp_shift<DPP_WAVE_SHL1>(v); }

__device__ __forceinline__ unsigned long long wave_max_u64m(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, MW);
        v = o > v ? o : v;
    }
    return v;
}

// Stage 1's wave-speed maxima of one block (its 4 waves) into the step slot.
// Every wave of the block calls this exactly once per tile, also a wave whose
// segment lies past the grid (with zeros): the maxima meet in LDS behind a
// barrier and one lane does the atomics.  Per-wave atomics on the slot's two
// addresses serialised when hundreds of short waves ended together (small
// grids); one per block is a quarter of them.  `par` alternates between
// consecutive tiles of a persistent block so the next tile's writes never
// meet this tile's reads.  The maxima are reached by a few blocks early on:
// the atomic is skipped when the slot already holds at least this value.
__device__ __forceinline__ void block_max_to_slot_m(unsigned long long smx, unsigned long long smy, StepSlot* s,
                                                    int par) {
    __shared__ unsigned long long red[2][MWAVES][2];
    smx = wave_max_u64m(smx);
    smy = wave_max_u64m(smy);
    const int lane = threadIdx.x & (MW - 1), w = threadIdx.x / MW;
    if (lane == 0) {
        red[par][w][0] = smx;
        red[par][w][1] = smy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 1; k < MWAVES; ++k) {
            smx = red[par][k][0] > smx ? red[par][k][0] : smx;
            smy = red[par][k][1] > smy ? red[par][k][1] : smy;
        }
        const unsigned long long cx = __hip_atomic_load(&s->wsmax[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long cy = __hip_atomic_load(&s->wsmax[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (smx > cx) atomicMax(&s->wsmax[0], smx);
        if (smy > cy) atomicMax(&s->wsmax[1], smy);
    }
}

__device__ __forceinline__ double Bload(const KParams& P, const double* __restrict__ B, int i, int j) {
    return (i <= P.nxT && j <= P.nyT) ? B[(size_t)i + (size_t)(P.nxT + 1) * j] : 0.0;
}

// Row access through a raw buffer resource: the descriptor (SGPRs) spans one
// row, base at its first element and `len` elements long, and a lane's
// access is its 32-bit byte offset (a loop-invariant VGPR).  That keeps the
// per-lane 64-bit address arithmetic of a global access out of the row loop,
// and the hardware range check reads zero for (and drops a store to) an
// offset past the row.  validate_config caps a row at 2^27 elements, so the
// byte offsets and lengths fit in 31 bits.  The row pointer must be the same for the whole wave.
constexpr int RSRC_WORD3 = 0x00020000;   // gfx9 raw buffer: 32-bit data format
template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const T* row, int len) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)row, (short)0, len * (int)sizeof(T), RSRC_WORD3);
}
__device__ __forceinline__ double rsrc_ld(__amdgpu_buffer_rsrc_t r, unsigned off, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float rsrc_ld(__amdgpu_buffer_rsrc_t r, unsigned off, float) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <class T>
__device__ __forceinline__ T row_ld(const T* row, int len, int col) {
    return rsrc_ld(row_rsrc(row, len), (unsigned)col * (unsigned)sizeof(T), T());
}
// Bload for a row j that is the same for the whole wave (the descriptor's
// length does Bload's range check: zero past column nxT or row nyT)
__device__ __forceinline__ double Bload_row(const KParams& P, const double* __restrict__ B, int i, int j) {
    return row_ld(B + (size_t)(P.nxT + 1) * j, j <= P.nyT ? P.nxT + 1 : 0, i);
}
__device__ __forceinline__ void row_st(double* row, int len, int col, double v) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(__amdgpu_buffer_rsrc_t(), 0u, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), row_rsrc(row, len), (unsigned)col * 8u, 0, 0);
}
__device__ __forceinline__ void row_st(float* row, int len, int col, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), row_rsrc(row, len), (unsigned)col * 4u, 0, 0);
}

// Stage-1 result of one cell (rungekutta.jl:38-41 with the Float64 dt,
// friction.jl:80-84 from U0 with the stored dt, timestepper.jl:53-63).
// FAST: the friction law is none or bh2021 (march_fast).
template <class S, bool FAST = false>
__device__ __forceinline__ void stage1_cell(const KParams& P, const S x[3], const S r[3], double Bm, double dt_d,
                                            S dt_s, S o[3]) {
    o[0] = (S)(x[0] + dt_d * r[0]);
    o[1] = (S)(x[1] + dt_d * r[1]);
    o[2] = (S)(x[2] + dt_d * r[2]);
    if (P.friction) {
        const double fs = friction_scalar(FAST ? 1 : P.friction, P.Cz, P.phys.kappa, x[0], x[1], x[2], Bm);
        apply_friction(o, dt_s, fs);
    }
    if (o[0] - Bm < P.phys.cutoff) {
        o[0] = (S)jl_max((double)o[0], Bm);
        o[1] = (S)0.0;
        o[2] = (S)0.0;
    }
}

// One wave's work: output columns gc + 60*strip .. +59 (lanes 2..61) of rows
// gc + seg*my .. +my-1.  Shared by the single-grid kernel (strip, seg from
// the block index) and the batched kernel (from the batch's tile table).
// Has no barriers: every wave is independent.
// FAST: the configuration is known to have linear reconstruction with a
// finite nonzero θ, the bottom source, the 2D cell-bottom lookup,
// power-of-two spacings, no friction or bh2021, no infiltration and a rain
// rate uniform in space (march_fast), so those run-time checks fold away.
template <class S, int STAGE, bool FAST = false, bool BATCH = false, bool RMP1 = false>
__device__ __forceinline__ void march_tile(const KParams& P, StepCtl* ctl, int slot, int my, int strip, int seg,
                                           const S* __restrict__ U0, const S* __restrict__ RHS1,
                                           S* __restrict__ OUT, const double* __restrict__ B, int par = 0,
                                           int close = 0, int prio_after = 0, int prio_total = 0,
                                           int prio_cap = 3) {
    const StepSlot& sl = ctl->slot[slot];
    const int lane = threadIdx.x & (MW - 1);
    if (!sl.active) {                              // the same for every wave of the block
        // an inactive step passes inactivity on, so steps launched past the
        // end of a run stay no-ops
        if (STAGE == 2 && close && strip == 0 && seg == 0 && lane == 0) ctl->slot[1 - slot].active = 0;
        return;
    }
    constexpr bool PREFETCH = STAGE == 1 ? SFV_MARCH_PREFETCH1 : SFV_MARCH_PREFETCH2;
    const int gc = P.gc, nxT = P.nxT, nyT = P.nyT;
    const int j0 = gc + seg * my;
    if (j0 >= nyT - gc) {                          // whole wave: past the grid
        if (STAGE == 1) block_max_to_slot_m(0ull, 0ull, &ctl->slot[slot], par);
        return;
    }
    const int th = min(my, nyT - gc - j0);
    const int i = gc + strip * MX + lane - 2;      // this lane's column (>= 0 since gc >= 2)
    const bool in_dom = i < nxT;
    const bool out_col = lane >= 2 && lane < MX + 2 && i < nxT - gc;
    const size_t n = (size_t)nxT * nyT;
    const bool rec_linear = FAST || P.rec == 1, bottom = FAST || P.bottom;
    auto bc4 = [&](double a, double b, double c, double d) {
        return FAST ? 0.25 * (((a + b) + c) + d) : bcell4(P, a, b, c, d);
    };
    auto sdiv = [&](double x, double dd, double inv, int pow2) { return FAST ? x * inv : slope_div(x, dd, inv, pow2); };
    // FAST: no infiltration and a rain rate that is the same for every cell
    auto sources = [&](S w, size_t k) -> S {
        if (FAST) return P.rain_kind ? (S)(w + sl.rain_now) : w;
        return apply_sources(P, ctl->slot[slot], w, k);
    };

    double dt_d = 0.0;
    S dt_s = S(0);
    if (STAGE == 2) {
        dt_d = stage1_dt<S>(P, ctl, sl);
        dt_s = (S)dt_d;
    }

    // Column remap of a ghost column to its boundary-condition source: stage
    // 2 always (it rebuilds U1 there from the source's U0 and RHS), stage 1
    // in its RMP1 instantiation (the ghost's value is its source's, with the
    // sign flip, as update_bc! would have stored it; corners, which the
    // boundary condition never writes, feed no output).  With both stages
    // reading through the maps, the march path never reads a ghost cell from
    // memory.
    constexpr bool rmp = STAGE == 2 || RMP1;
    // Only the edge strips and segments hold ghost columns or rows; the test
    // is the same for the whole wave, so interior waves skip the maps.  The
    // lanes span columns gc + strip*MX - 2 .. gc + strip*MX + 61, the rows
    // loaded are j0 - 2 .. j0 + th + 1.
    const bool xedge = rmp && (strip == 0 || gc + strip * MX + MW - 3 >= nxT - gc);
    const bool yedge = rmp && (j0 - 2 < gc || j0 + th + 1 >= nyT - gc);
    int sci = i, fxc = 0;
    if (xedge && in_dom) {
        if (i < gc) { sci = P.src_lo_x[i]; fxc = P.flip_lo_x[i]; }
        else if (i >= nxT - gc) { sci = P.src_hi_x[i - (nxT - gc)]; fxc = P.flip_hi_x[i - (nxT - gc)]; }
    }

    // ---- raw inputs, fetched one row ahead so the loads overlap a row of work
    struct Raw {
        S x[3];          // stage input (stage 1) or U0 (stage 2) at the source cell
        S r[3];          // RHS1 at the source cell (stage 2)
        double bsrc;     // B_cell of a remapped source cell (stage 2)
        int fy;          // hv negated (stage 2, ghost row of a wall)
        bool remap;
    };
    // Every lane loads: a lane past the grid (i >= nxT, so sci = i) reads
    // zeros through the row descriptors' range check.  No branch surrounds
    // the loads, so the loads of several rows can be in flight together
    // (a load under a partial exec mask must complete before other lanes
    // write its registers).
    auto fetch = [&](int jr, Raw& o) {
        o.remap = false;
        o.fy = 0;
        o.bsrc = 0.0;
        int sj = jr;
        if (yedge) {
            if (jr < gc) { sj = P.src_lo_y[jr]; o.fy = P.flip_lo_y[jr]; }
            else if (jr >= nyT - gc) { sj = P.src_hi_y[jr - (nyT - gc)]; o.fy = P.flip_hi_y[jr - (nyT - gc)]; }
        }
        const size_t ro = (size_t)nxT * sj;       // row sj (uniform), column sci (per lane)
        o.x[0] = row_ld(U0 + ro, nxT, sci); o.x[1] = row_ld(U0 + n + ro, nxT, sci);
        o.x[2] = row_ld(U0 + 2 * n + ro, nxT, sci);
        if (STAGE == 1) {
            // a lane past the grid may take -0.0 here; it feeds no output
            // and no recorded face
            if (xedge && fxc) o.x[1] = (S)(-o.x[1]);
            if (o.fy) o.x[2] = (S)(-o.x[2]);
        }
        if (STAGE == 2) {
            o.r[0] = row_ld(RHS1 + ro, nxT, sci); o.r[1] = row_ld(RHS1 + n + ro, nxT, sci);
            o.r[2] = row_ld(RHS1 + 2 * n + ro, nxT, sci);
            o.remap = in_dom && ((xedge && sci != i) || sj != jr);
            if (o.remap)   // the source cell's bottom
                o.bsrc = bc4(Bload_row(P, B, sci, sj), Bload_row(P, B, sci + 1, sj), Bload_row(P, B, sci, sj + 1),
                             Bload_row(P, B, sci + 1, sj + 1));
        }
    };
    // bottom corner row: own column, and (lane 63 only) the next column
    auto fetch_b = [&](int row, double& own, double& extra) {
        const double a = Bload_row(P, B, i, row), b = Bload_row(P, B, i + 1, row);   // all lanes, as above
        own = in_dom ? a : 0.0;
        extra = (lane == MW - 1 && in_dom) ? b : 0.0;
    };
    auto next_col = [&](double own, double extra) {
        const double v = shfl_dn1(own);
        return lane == MW - 1 ? extra : v;
    };

    // corner window: rows jr-1 (bm1), jr (b0), jr+1 (bp1); n* = column i+1
    double bm1 = 0.0, nm1 = 0.0, b0, n0, bp1, np1, e0_, e1_;
    fetch_b(j0 - 2, b0, e0_);
    n0 = next_col(b0, e0_);
    fetch_b(j0 - 1, bp1, e1_);
    np1 = next_col(bp1, e1_);
    Raw cur, nxt;
    if (PREFETCH) fetch(j0 - 2, cur);
    // The first NPRE rows' inputs and the corner rows read with them are
    // loaded together, so a wave waits for memory once before its first row
    // rather than once per row (these loads are the ones the peeled rows
    // r < NPRE would issue: rows j0-2+r and corner rows j0+r, r <= 3).
    constexpr int NPRE = PREFETCH ? 0 : SFV_MARCH_STARTFETCH;
    Raw pre[NPRE > 0 ? NPRE : 1];
    double preb[NPRE > 0 ? NPRE : 1], pree[NPRE > 0 ? NPRE : 1];
#pragma unroll
    for (int k = 0; k < NPRE; ++k) {
        fetch(j0 - 2 + k, pre[k]);
        fetch_b(j0 + k, preb[k], pree[k]);
    }
    // ranged face evaluation (central_upwind_rng) while the wave's inputs allow it
    bool band = P.ranged && __all(in_band5(b0, e0_, bp1, e1_, 0.0));

    S cA[3] = {S(0), S(0), S(0)}, cB[3] = {S(0), S(0), S(0)};
    // stage 2 with SFV_MARCH_KEEPU0: U0 of rows jr-2, jr-1 as loaded, for the
    // RK combination of row jr-2 (instead of loading it again)
    S xA[3] = {S(0), S(0), S(0)}, xB[3] = {S(0), S(0), S(0)};
    S Lyp0 = S(0), Ryp[3] = {S(0), S(0), S(0)};   // previous row's y faces (left: depth only)
    double Blyp = 0.0, Bryp = 0.0;                 // its y face bottoms
    double Fyp[3] = {0.0, 0.0, 0.0};               // y flux below that row
    S accx[3] = {S(0), S(0), S(0)};                // x part of the pending row
    double bcp = 0.0;                              // B_cell of the pending row
    unsigned long long smx = 0, smy = 0;
    const double g = P.phys.g;
    // x faces this lane evaluates (its left face) that border an output cell
    const bool face_x_ok = in_dom && lane >= 2 && lane <= MX + 2 && (lane == 2 ? i < nxT - gc : i - 1 < nxT - gc);

    // Work on row q = jr-1 is done only where it feeds an output row: its x
    // part (x reconstruction, x faces, x flux difference and bottom term) when
    // q is an output row, j0 <= q < j0+th, i.e. 3 <= r <= th+2; the y face
    // below it when it borders one, j0-1 < q <= j0+th, i.e. r >= 3.  The row
    // loop is r = 0..th+3: the first four iterations and the last are peeled,
    // so the loop proper (4 <= r <= th+2) does everything in one basic block.
    auto body = [&](int r, auto full_tag, auto x_tag) {
        constexpr bool FULL = decltype(full_tag)::value;   // r >= 4: every step applies
        const int jr = j0 - 2 + r;
#if SFV_MARCH_TIMELINE
        if (!BATCH && r < TL_ROWS && lane == 0) {
            const int tw = blockIdx.x * MWAVES + threadIdx.x / MW;
            if (tw < TL_WAVES) tl_rows[STAGE - 1][tw][r] = __builtin_amdgcn_s_memrealtime();
        }
#endif
        // Issue priority by iterations left (this segment's and, in the
        // persistent batch kernel, those of the wave's later tiles): 3 until
        // the last 3/DIV of the wave's total, then one step lower every
        // 1/DIV.  The waves sharing a SIMD then progress together and end
        // together, instead of the oldest finishing first and the youngest
        // running on alone at the end of the launch.  prio_after < 0: the
        // caller keeps its own priority; prio_cap: the highest level used.
        if (SFV_MARCH_PRIO && prio_after >= 0) {
            const int left = th + 3 - r + prio_after,
                      q = max(1, (prio_total > 0 ? prio_total : th + 4) / SFV_MARCH_PRIO_DIV);
            set_issue_priority(min(prio_cap, left >= 3 * q ? 3 : left >= 2 * q ? 2 : left >= q ? 1 : 0));
        }
        const bool xrow = FULL ? decltype(x_tag)::value : r == 3;
        const bool yface = FULL || r >= 3;
        double nb, ne;
        const bool pre_r = !FULL && r < NPRE;      // a peeled row (r constant) loaded up front
        if (pre_r) {
            nb = preb[r];
            ne = pree[r];
        } else {
            fetch_b(jr + 2, nb, ne);               // next corner row (Bload bounds-checks row nyT+1)
        }
        if (PREFETCH) {
            if (r + 1 < th + 4) fetch(jr + 1, nxt);    // rows up to j0+th+1 <= nyT-1 (gc >= 2)
        } else if (pre_r) {
            cur = pre[r];
        } else {
            fetch(jr, cur);
        }

        // stage input of row jr
        S cC[3];
        if (STAGE == 1) {
            cC[0] = cur.x[0]; cC[1] = cur.x[1]; cC[2] = cur.x[2];
        } else if (in_dom) {
            const double Bm = cur.remap ? cur.bsrc : bc4(b0, n0, bp1, np1);
            stage1_cell<S, FAST>(P, cur.x, cur.r, Bm, dt_d, dt_s, cC);
            if (fxc) cC[1] = (S)(-cC[1]);
            if (cur.fy) cC[2] = (S)(-cC[2]);
        } else {
            cC[0] = cC[1] = cC[2] = S(0);
        }
        // the ranged face forms apply while every stage input and bottom
        // corner this wave has taken in lies in band (central_upwind_rng);
        // the faces of this iteration read rows up to jr and corner rows up
        // to jr, the corner row jr+2 taken in here is read two rows later
        band = band && __all(in_band5(cC[0], cC[1], cC[2], nb, ne));

        // 1. y reconstruction of row m = jr-1 (rows jr-2, jr-1, jr) and x
        //    reconstruction of row q = jr-1 (this lane's cell cB and its x
        //    neighbours; bottom corner rows jr-1, jr)
        S Ly[3] = {S(0), S(0), S(0)}, Ry[3] = {S(0), S(0), S(0)};
        const double Bly = 0.5 * (bm1 + nm1), Bry = 0.5 * (b0 + n0);
        const double bcq = bc4(bm1, nm1, b0, n0);                 // B_cell of row jr-1
        S Lx[3] = {S(0), S(0), S(0)}, Rx[3] = {S(0), S(0), S(0)}, Rxl[3] = {S(0), S(0), S(0)};
        const double Blx = 0.5 * (bm1 + b0), Brx = 0.5 * (nm1 + n0);
        if (FULL || r >= 2) {
            if (rec_linear) {
#if SFV_MARCH_BRANCHFREE
                reconstruct_cell_nb(cA, cB, cC, P.theta, Bly, Bry, Ly, Ry);
#else
                reconstruct_cell(cA, cB, cC, P.theta, Bly, Bry, Ly, Ry, FAST);
#endif
            } else {
#pragma unroll
                for (int v = 0; v < 3; ++v) { Ly[v] = Ry[v] = cB[v]; }
                Ly[0] = Ry[0] = (S)(cB[0] - bcq);
            }
        }
        if (xrow) {
            if (rec_linear) {
                S cl[3], cr[3];
#pragma unroll
                for (int v = 0; v < 3; ++v) { cl[v] = shfl_up1(cB[v]); cr[v] = shfl_dn1(cB[v]); }
#if SFV_MARCH_BRANCHFREE
                reconstruct_cell_nb(cl, cB, cr, P.theta, Blx, Brx, Lx, Rx);
#else
                reconstruct_cell(cl, cB, cr, P.theta, Blx, Brx, Lx, Rx, FAST);
#endif
            } else {
#pragma unroll
                for (int v = 0; v < 3; ++v) { Lx[v] = Rx[v] = cB[v]; }
                Lx[0] = Rx[0] = (S)(cB[0] - bcq);
            }
#pragma unroll
            for (int v = 0; v < 3; ++v) Rxl[v] = shfl_up1(Rx[v]);
        }
        // 2. the faces: the y face between rows jr-2 and jr-1, and this lane's
        //    left x face in row jr-1 (independent chains, one basic block when
        //    both apply)
        double Fy[3] = {0.0, 0.0, 0.0}, Fxm[3] = {0.0, 0.0, 0.0};
        auto face_y = [&](auto rng_tag) {
#if SFV_MARCH_WETFAST
            const double spy = central_upwind_auto<1>(Ryp, Ly, P.phys, Fy);
#else
            const double spy = decltype(rng_tag)::value ? central_upwind_rng(1, Ryp, Ly, P.phys, Fy)
                                                        : central_upwind(1, Ryp, Ly, P.phys, Fy);
#endif
            if (STAGE == 1) {
                const int m = jr - 1;
                if (out_col && m >= j0 && m <= j0 + th) {
                    const unsigned long long bits = (unsigned long long)__double_as_longlong((double)(S)spy);
                    smy = bits > smy ? bits : smy;
                }
            }
        };
        auto face_x = [&](auto rng_tag) {
#if SFV_MARCH_WETFAST
            const double spx = central_upwind_auto<0>(Rxl, Lx, P.phys, Fxm);
#else
            const double spx = decltype(rng_tag)::value ? central_upwind_rng(0, Rxl, Lx, P.phys, Fxm)
                                                        : central_upwind(0, Rxl, Lx, P.phys, Fxm);
#endif
            if (STAGE == 1 && face_x_ok) {       // xrow: row jr-1 is an output row
                const unsigned long long bits = (unsigned long long)__double_as_longlong((double)(S)spx);
                smx = bits > smx ? bits : smx;
            }
        };
#if SFV_MARCH_BRANCHFREE
        if (yface || xrow) {
            const S* hp[4] = {&Ryp[0], &Ly[0], &Rxl[0], &Lx[0]};
            double hs[4];
            desingularize_n<S, 4>(hp, P.phys.kappa, hs);
            if (yface) {
                const double spy = central_upwind_nb<1>(Ryp, Ly, hs[0], hs[1], P.phys, Fy);
                if (STAGE == 1 && out_col && jr - 1 >= j0 && jr - 1 <= j0 + th) {
                    const unsigned long long bits = (unsigned long long)__double_as_longlong((double)(S)spy);
                    smy = bits > smy ? bits : smy;
                }
            }
            if (xrow) {
                const double spx = central_upwind_nb<0>(Rxl, Lx, hs[2], hs[3], P.phys, Fxm);
                if (STAGE == 1 && face_x_ok) {
                    const unsigned long long bits = (unsigned long long)__double_as_longlong((double)(S)spx);
                    smx = bits > smx ? bits : smx;
                }
            }
        }
#else
        if (band) {                                // every input of this row's faces is in band
            if (yface) face_y(std::true_type{});
            if (xrow) face_x(std::true_type{});
        } else {
            if (yface) face_y(std::false_type{});
            if (xrow) face_x(std::false_type{});
        }
#endif
        // 3. x part of row jr-1 (finished next iteration)
        S accx_n[3] = {S(0), S(0), S(0)};
        if (xrow) {
            double Fxp[3];
#pragma unroll
            for (int v = 0; v < 3; ++v) Fxp[v] = shfl_dn1(Fxm[v]);
#pragma unroll
            for (int v = 0; v < 3; ++v) accx_n[v] = (S)(S(0) - P.inv_dx * (Fxp[v] - Fxm[v]));
            if (bottom) {
                const S hsum = Rx[0] + Lx[0];
                accx_n[1] = (S)(accx_n[1] + -g * sdiv(Brx - Blx, P.dx, P.inv_dx, P.dx_pow2) * (hsum / 2.0));
            }
        }
        // 4. finish row p = jr-2: y part on top of its x part, then sources
        if ((FULL || r >= 4) && out_col) {
            const int p = jr - 2;
            const size_t k = (size_t)i + (size_t)nxT * p;     // for the sources' per-cell tables
            S* orow = OUT + (size_t)nxT * p;                   // row p (uniform), column i (per lane)
            S acc[3] = {accx[0], accx[1], accx[2]};
#pragma unroll
            for (int v = 0; v < 3; ++v) acc[v] = (S)(acc[v] - P.inv_dy * (Fy[v] - Fyp[v]));
            if (bottom) {
                const S hsum = Ryp[0] + Lyp0;
                acc[2] = (S)(acc[2] + -g * sdiv(Bryp - Blyp, P.dy, P.inv_dy, P.dy_pow2) * (hsum / 2.0));
            }
            acc[0] = sources(acc[0], k);
            if (STAGE == 1) {
                row_st(orow, nxT, i, acc[0]);
                row_st(orow + n, nxT, i, acc[1]);
                row_st(orow + 2 * n, nxT, i, acc[2]);
            } else {
                // cA is U1 of row p
                S o[3];
#pragma unroll
                for (int v = 0; v < 3; ++v) o[v] = (S)(cA[v] + dt_s * acc[v]);
                if (SFV_MARCH_KEEPU0) {
                    // output rows are interior, so xA holds U0 at (i, p) itself
                    o[0] = (S)(0.5 * (xA[0] + o[0]));
                    o[1] = (S)(0.5 * (xA[1] + o[1]));
                    o[2] = (S)(0.5 * (xA[2] + o[2]));
                } else {
                    o[0] = (S)(0.5 * (U0[k] + o[0]));
                    o[1] = (S)(0.5 * (U0[n + k] + o[1]));
                    o[2] = (S)(0.5 * (U0[2 * n + k] + o[2]));
                }
                if (P.friction) {
                    const double fs = friction_scalar(FAST ? 1 : P.friction, P.Cz, P.phys.kappa, cA[0], cA[1], cA[2], bcp);
                    apply_friction(o, dt_s, fs);
                }
                const bool dry = o[0] - bcp < P.phys.cutoff;
                const S wd = (S)jl_max((double)o[0], bcp);
                o[0] = dry ? wd : o[0];
                o[1] = dry ? (S)0.0 : o[1];
                o[2] = dry ? (S)0.0 : o[2];
                row_st(orow, nxT, i, o[0]);
                row_st(orow + n, nxT, i, o[1]);
                row_st(orow + 2 * n, nxT, i, o[2]);
            }
        }
        // shift the windows
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            cA[v] = cB[v]; cB[v] = cC[v];
            if (STAGE == 2 && SFV_MARCH_KEEPU0) { xA[v] = xB[v]; xB[v] = cur.x[v]; }
            Ryp[v] = Ry[v];
            Fyp[v] = Fy[v];
            accx[v] = accx_n[v];
        }
        Lyp0 = Ly[0];
        Blyp = Bly; Bryp = Bry;
        bcp = bcq;
        bm1 = b0; nm1 = n0;
        b0 = bp1; n0 = np1;
        bp1 = nb; np1 = next_col(nb, ne);
        if (PREFETCH) cur = nxt;
    };

    // th >= 1, so there are th+4 >= 5 iterations: r = 0..3 peeled, the loop
    // proper r = 4..th+2 (x part and y face), and r = th+3 (y face only).
    body(0, std::false_type{}, std::false_type{});
    body(1, std::false_type{}, std::false_type{});
    body(2, std::false_type{}, std::false_type{});
    body(3, std::false_type{}, std::false_type{});
    // The row loop is unrolled by hand (its cross-lane shifts are convergent
    // operations, which the compiler will not unroll with a runtime trip count).
    int r = 4;
    constexpr int UNROLL = STAGE == 1 ? SFV_MARCH_UNROLL1 : SFV_MARCH_UNROLL2;
    if (UNROLL > 1)
        for (; r + UNROLL <= th + 3; r += UNROLL) unrolled_rows(body, r, std::make_index_sequence<UNROLL>{});
    for (; r < th + 3; ++r) body(r, std::true_type{}, std::true_type{});
    body(th + 3, std::true_type{}, std::false_type{});

    if (STAGE == 1) block_max_to_slot_m(smx, smy, &ctl->slot[slot], par);

    // dt for the host and the step's closing (t += dt, next slot armed), by
    // one lane: the other waves read only this step's slot, and the next
    // slot is read only by the next launches.
    if (STAGE == 2 && strip == 0 && seg == 0 && lane == 0) {
        ctl->dt = (double)dt_s;
        ctl->dt_stage1 = dt_d;
        if (close) {
            if (BATCH) close_step_call<S, FAST>(P, ctl, slot);
            else close_step<S, FAST>(P, ctl, slot);
        }
    }

    // The gc x gc corner cells of the domain: never written by the boundary
    // condition (bc.jl:92-109) nor read by a stencil, but part of the state;
    // both stages run sources, RK update, friction and clamp on them.
    if (STAGE == 2 && strip == 0 && seg == 0) {
        for (int e = lane; e < 4 * gc * gc; e += MW) {
            const int corner = e / (gc * gc), rr = e % (gc * gc);
            const int ci = (corner & 1) ? nxT - gc + rr % gc : rr % gc;
            const int cj = (corner >> 1) ? nyT - gc + rr / gc : rr / gc;
            const size_t k = (size_t)ci + (size_t)nxT * cj;
            const double Bm = bc4(Bload(P, B, ci, cj), Bload(P, B, ci + 1, cj), Bload(P, B, ci, cj + 1),
                                     Bload(P, B, ci + 1, cj + 1));
            const S src = sources(S(0), k);
            const S x[3] = {U0[k], U0[n + k], U0[2 * n + k]};
            const S rs[3] = {src, S(0), S(0)};
            S u1[3];
            stage1_cell<S, FAST>(P, x, rs, Bm, dt_d, dt_s, u1);
            S o[3];
#pragma unroll
            for (int v = 0; v < 3; ++v) o[v] = (S)(u1[v] + dt_s * rs[v]);
#pragma unroll
            for (int v = 0; v < 3; ++v) o[v] = (S)(0.5 * (x[v] + o[v]));
            if (P.friction) {
                const double fs = friction_scalar(FAST ? 1 : P.friction, P.Cz, P.phys.kappa, u1[0], u1[1], u1[2], Bm);
                apply_friction(o, dt_s, fs);
            }
            if (o[0] - Bm < P.phys.cutoff) {
                o[0] = (S)jl_max((double)o[0], Bm);
                o[1] = (S)0.0;
                o[2] = (S)0.0;
            }
            OUT[k] = o[0];
            OUT[n + k] = o[1];
            OUT[2 * n + k] = o[2];
        }
    }
}

// The hardware hands workgroup L of a launch to XCD L mod 8, and each XCD
// has its own L2.  Block L therefore takes logical tile
//   prefix(L mod 8) + L div 8,
// so the blocks of one XCD work on one contiguous run of tiles, where
// neighbouring segments (which share 4 halo rows) are resident together.
constexpr int N_XCD = 8;
__device__ __forceinline__ int xcd_tile(int L, int N) {
#if SFV_MARCH_XCD
    const int x = L % N_XCD, k = L / N_XCD, q = N / N_XCD, r = N % N_XCD;
    return x * q + min(x, r) + k;
#else
    (void)N;
    return L;
#endif
}

// Does the wave of (strip, seg) load a ghost column or row?  Its lanes span
// columns gc + strip*MX - 2 .. gc + strip*MX + 61 and it loads rows
// j0 - 2 .. j0 + th + 1.  Only such waves run stage 1's remapping code.
__device__ __forceinline__ bool march_edge(const KParams& P, int my, int strip, int seg) {
    const int gc = P.gc, j0 = gc + seg * my;
    const int th = min(my, P.nyT - gc - j0);
    return strip == 0 || gc + strip * MX + MW - 3 >= P.nxT - gc || j0 - 2 < gc || j0 + th + 1 >= P.nyT - gc;
}

// Grid of march_blocks(strips, segs) blocks of MWAVES waves (swe_kernels.hpp).
// With SFV_MARCH_LINEAR a wave's number is the block's logical tile times
// MWAVES plus the wave, and a workgroup's waves are neighbouring strips of
// one segment row.
template <class S, int STAGE, bool FAST>
__global__ void __launch_bounds__(MW* MWAVES)
__attribute__((amdgpu_waves_per_eu(STAGE == 1 ? SFV_MARCH_MIN_WAVES1 : SFV_MARCH_MIN_WAVES)))
k_march(KParams P, StepCtl* ctl, int slot, int my, int strips, int segs, const S* __restrict__ U0,
        const S* __restrict__ RHS1, S* __restrict__ OUT, const double* __restrict__ B, int close, int remap) {
    const int t = xcd_tile(blockIdx.x, march_blocks(strips, segs));
    // the same for every lane of the wave; readfirstlane tells the compiler so
#if SFV_MARCH_LINEAR
    const int w = __builtin_amdgcn_readfirstlane(t * MWAVES + threadIdx.x / MW);
    const bool idle = w >= strips * segs;          // past the last tile: seg = segs is past the grid
#if SFV_MARCH_XCD == 2
    const int strip = idle ? 0 : w % strips, seg = idle ? segs : w / strips;
#else
    const int strip = idle ? 0 : w / segs, seg = idle ? segs : w % segs;
#endif
#else
    const int seggroups = (segs + MWAVES - 1) / MWAVES;
#if SFV_MARCH_XCD == 2
    const int strip = t % strips, sg = t / strips;
#else
    const int strip = t / seggroups, sg = t % seggroups;
#endif
    const int seg = __builtin_amdgcn_readfirstlane(sg * MWAVES + threadIdx.x / MW);
#endif
#if SFV_MARCH_TIMELINE
    const unsigned long long tl0 = __builtin_amdgcn_s_memrealtime();
#endif
    if (STAGE == 1 && remap && march_edge(P, my, strip, seg))
        march_tile<S, 1, FAST, false, true>(P, ctl, slot, my, strip, seg, U0, RHS1, OUT, B);
    else
        march_tile<S, STAGE, FAST>(P, ctl, slot, my, strip, seg, U0, RHS1, OUT, B, 0, close);
#if SFV_MARCH_TIMELINE
    const unsigned long long tl1 = __builtin_amdgcn_s_memrealtime();
    const int tw = blockIdx.x * MWAVES + threadIdx.x / MW;
    if ((threadIdx.x & (MW - 1)) == 0 && tw < TL_WAVES) {
        unsigned long long* o = tl_wave[STAGE - 1][tw];
        o[0] = tl0;
        o[1] = tl1;
        o[2] = __builtin_amdgcn_s_getreg((31 << 11) | 4);                       // HW_ID
        o[3] = ((unsigned long long)__builtin_amdgcn_s_getreg((15 << 11) | 20) << 32) |   // XCC_ID
               ((unsigned long long)(strip & 0xFFFF) << 16) | (unsigned)(seg & 0xFFFF);
    }
#endif
}

// Batched stage kernel: a persistent grid (a multiple of 8 workgroups) walks
// the tiles of every member.  The tile order is split into 8 contiguous
// chunks, one per XCD, and the workgroups of XCD x (blockIdx mod 8 == x) walk
// chunk x, so a member's neighbouring tiles share that XCD's L2.  With
// hdr->prio == 1 (a batch of small grids stepping beside another batch) the
// waves take the highest issue priority: their short serial chains issue
// first and the other batch's waves fill the gaps; that other batch runs with
// hdr->prio == 2, its rows-left priority capped one level below.
template <class S, int STAGE, bool FAST>
__global__ void __launch_bounds__(MW* MWAVES) __attribute__((amdgpu_waves_per_eu(SFV_MARCH_MIN_WAVES)))
k_march_batch(const BatchMember* __restrict__ mem, const BatchHeader* __restrict__ hdr, int lp) {
    const int members = hdr->members, tiles = hdr->tiles;
    if (hdr->prio == 1) __builtin_amdgcn_s_setprio(3);
    const int cap = hdr->prio == 2 ? 2 : 3;          // a host engine's waves stay below its guest's
    const int x = blockIdx.x % N_XCD, k = blockIdx.x / N_XCD, per = gridDim.x / N_XCD;
    const int q = tiles / N_XCD, r = tiles % N_XCD;
    const int c0 = x * q + min(x, r), c1 = c0 + q + (x < r ? 1 : 0);
    int par = 0;
    // this wave's tiles: c0 + k, c0 + k + per, ... below c1
    const int ntiles = c1 > c0 + k ? (c1 - 1 - c0 - k) / per + 1 : 0;
    for (int t = c0 + k; t < c1; t += per, par ^= 1) {
        const int m = batch_find(mem, members, t);
        const BatchMember& M = mem[m];
        const int local = t - M.tile0;
        const int cur = (M.cur0 ^ lp) & 1, slot = (M.slot0 ^ lp) & 1;
        // the same for every lane of the wave; readfirstlane tells the compiler so
#if SFV_MARCH_LINEAR
        const int w = __builtin_amdgcn_readfirstlane(local * MWAVES + threadIdx.x / MW);
        const bool idle = w >= M.strips * M.segs;  // past the member's last tile: seg = segs is past its grid
        const int strip = idle ? 0 : w % M.strips, seg = idle ? M.segs : w / M.strips;
#else
        const int strip = local % M.strips, sg = local / M.strips;
        const int seg = __builtin_amdgcn_readfirstlane(sg * MWAVES + threadIdx.x / MW);
#endif
        const S* U0 = (const S*)M.U[cur];
        // iterations of this wave's later tiles, counted at this member's height
        const int after = hdr->prio == 1 ? -1 : ((c1 - 1 - t) / per) * (M.my + 4);
        const int total = ntiles * (M.my + 4);
        if (STAGE == 1)
            if (march_edge(M.P, M.my, strip, seg))
                march_tile<S, 1, FAST, true, true>(M.P, M.ctl, slot, M.my, strip, seg, U0, nullptr, (S*)M.RHS, M.B, par,
                                                   0, after, total, cap);
            else
                march_tile<S, 1, FAST, true>(M.P, M.ctl, slot, M.my, strip, seg, U0, nullptr, (S*)M.RHS, M.B, par, 0,
                                             after, total, cap);
        else
            march_tile<S, 2, FAST, true>(M.P, M.ctl, slot, M.my, strip, seg, U0, (const S*)M.RHS, (S*)M.U[cur ^ 1], M.B, 0,
                                         1, after, total, cap);
    }
}

// The configurations the FAST kernels cover.
bool march_fast(const KParams& P) {
    return SFV_MARCH_FAST && P.rec == 1 && std::isfinite(P.theta) && P.theta != 0.0 && P.bottom && !P.bcell_1d &&
           P.dx_pow2 && P.dy_pow2 && (P.friction == 0 || P.friction == 1) && P.infil_kind == 0 &&
           (P.rain_kind == 0 || P.rain_kind == 1 || P.rain_kind == 2);
}

// Rows per wave segment for this grid.  A launch runs in rounds of resident
// workgroups (CUs x workgroups per CU), and each round lasts about one
// segment's march (its rows + 4 halo rows), so the makespan is about
//   rounds x (my + 4),  rounds = ceil(march_blocks(strips, segments) / resident);
// the height in [1, 64] that minimises it wins (the smallest on a tie).  At
// 2048^2 that is 24 rows: 35 x 86 waves in 753 workgroups, one round of the
// 768 resident ones (23 rows would need 788); at 512^2, 2 rows.
template <class S, int STAGE, bool FAST>
static int march_resident_blocks() {
    int blocks = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_march<S, STAGE, FAST>, MW * MWAVES, 0);
    return (e == hipSuccess && blocks > 0) ? blocks : 2;
}
int march_rows_per_wave(const KParams& P, int stage, int precision) {
    if (SFV_MARCH_ROWS > 0) return SFV_MARCH_ROWS;
    static int cus = 0, bpc[2][2][2];       // [precision][stage - 1][fast]
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        if (cus <= 0) cus = 256;
        bpc[0][0][0] = march_resident_blocks<double, 1, false>();
        bpc[0][0][1] = march_resident_blocks<double, 1, true>();
        bpc[0][1][0] = march_resident_blocks<double, 2, false>();
        bpc[0][1][1] = march_resident_blocks<double, 2, true>();
        bpc[1][0][0] = march_resident_blocks<float, 1, false>();
        bpc[1][0][1] = march_resident_blocks<float, 1, true>();
        bpc[1][1][0] = march_resident_blocks<float, 2, false>();
        bpc[1][1][1] = march_resident_blocks<float, 2, true>();
    });
    const long resident = (long)cus * bpc[precision ? 1 : 0][stage == 1 ? 0 : 1][march_fast(P) ? 1 : 0];
    const long strips = (P.nx + MX - 1) / MX;
    int best = 1;
    long best_cost = -1;
    for (int my = 1; my <= 64; ++my) {
        const long blocks = march_blocks((int)strips, (P.ny + my - 1) / my);
        const long rounds = (blocks + resident - 1) / resident;
        const long cost = rounds * (my + 4);
        if (best_cost < 0 || cost < best_cost) { best_cost = cost; best = my; }
    }
    return best;
}

template <class S>
hipError_t launch_march(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, int stage, const S* U0,
                        const S* RHS1, S* OUT, const double* B, int close, int remap_halo) {
    const int my = P.rows > 0 ? P.rows : march_rows_per_wave(P, stage, std::is_same<S, float>::value ? 1 : 0);
    const int segs = (P.ny + my - 1) / my;
    const int strips = (P.nx + MX - 1) / MX;
    const dim3 grid(march_blocks(strips, segs));
    const dim3 block(MW * MWAVES);
    if (march_fast(P)) {
        if (stage == 1)
            hipLaunchKernelGGL((k_march<S, 1, true>), grid, block, 0, st, P, ctl, slot, my, strips, segs, U0, RHS1,
                               OUT, B, close, remap_halo);
        else
            hipLaunchKernelGGL((k_march<S, 2, true>), grid, block, 0, st, P, ctl, slot, my, strips, segs, U0, RHS1,
                               OUT, B, close, remap_halo);
    } else {
        if (stage == 1)
            hipLaunchKernelGGL((k_march<S, 1, false>), grid, block, 0, st, P, ctl, slot, my, strips, segs, U0,
                               RHS1, OUT, B, close, remap_halo);
        else
            hipLaunchKernelGGL((k_march<S, 2, false>), grid, block, 0, st, P, ctl, slot, my, strips, segs, U0,
                               RHS1, OUT, B, close, remap_halo);
    }
    return hipPeekAtLastError();
}
template <class S>
hipError_t launch_march_batch(hipStream_t st, int stage, const BatchMember* mem, const BatchHeader* hdr, int lp,
                              int blocks, bool fast) {
    if (blocks < N_XCD || blocks % N_XCD) return hipErrorInvalidValue;
    const dim3 grid(blocks), block(MW * MWAVES);
    if (fast) {
        if (stage == 1) hipLaunchKernelGGL((k_march_batch<S, 1, true>), grid, block, 0, st, mem, hdr, lp);
        else hipLaunchKernelGGL((k_march_batch<S, 2, true>), grid, block, 0, st, mem, hdr, lp);
    } else {
        if (stage == 1) hipLaunchKernelGGL((k_march_batch<S, 1, false>), grid, block, 0, st, mem, hdr, lp);
        else hipLaunchKernelGGL((k_march_batch<S, 2, false>), grid, block, 0, st, mem, hdr, lp);
    }
    return hipPeekAtLastError();
}

int march_batch_blocks_per_cu(int precision, int stage) {
    int blocks = 0;
    hipError_t e;
    if (precision == 0)
        e = stage == 1
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_march_batch<double, 1, false>, MW * MWAVES, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_march_batch<double, 2, false>, MW * MWAVES, 0);
    else
        e = stage == 1
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_march_batch<float, 1, false>, MW * MWAVES, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_march_batch<float, 2, false>, MW * MWAVES, 0);
    return (e == hipSuccess && blocks > 0) ? blocks : 2;
}

template hipError_t launch_march_batch<double>(hipStream_t, int, const BatchMember*, const BatchHeader*, int, int,
                                               bool);
template hipError_t launch_march_batch<float>(hipStream_t, int, const BatchMember*, const BatchHeader*, int, int,
                                              bool);
template hipError_t launch_march<double>(hipStream_t, const KParams&, StepCtl*, int, int, const double*,
                                         const double*, double*, const double*, int, int);
template hipError_t launch_march<float>(hipStream_t, const KParams&, StepCtl*, int, int, const float*, const float*,
                                        float*, const double*, int, int);

}  // namespace sfv

#if SFV_MARCH_TIMELINE
// Measurement build only: copy the timeline buffers out (waves: 2 x 16384 x 4,
// rows: 2 x 16384 x 32 unsigned 64-bit words) or clear them.
extern "C" int sfv_probe_timeline(unsigned long long* waves, unsigned long long* rows) {
    hipError_t e = hipMemcpyFromSymbol(waves, HIP_SYMBOL(sfv::tl_wave), sizeof(sfv::tl_wave), 0, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpyFromSymbol(rows, HIP_SYMBOL(sfv::tl_rows), sizeof(sfv::tl_rows), 0, hipMemcpyDeviceToHost);
    return (int)e;
}
extern "C" int sfv_probe_timeline_clear() {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(sfv::tl_wave));
    if (e == hipSuccess) e = hipMemset(p, 0, sizeof(sfv::tl_wave));
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(sfv::tl_rows));
    if (e == hipSuccess) e = hipMemset(p, 0, sizeof(sfv::tl_rows));
    return (int)e;
}
#endif
