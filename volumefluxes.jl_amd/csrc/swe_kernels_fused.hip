// swe_kernels_fused.hip — one kernel per RK2 stage, tiled through LDS.
//
// A block owns a FT_X x FT_Y tile of interior cells.  It stages the stage
// input on the tile plus a 2-cell halo (the reconstruction stencil) and the
// bottom corners in LDS with coalesced x-fastest loads, evaluates every face
// flux of the tile ONCE into an LDS face buffer (the reference evaluates each
// face from both sides, numericalflux.jl:25-26, with identical arguments, so
// sharing is bit-exact), then updates each cell from its two faces per
// direction.  Faces on a tile edge are evaluated by both neighbouring blocks.
//
// k_stage1: U0 -> RHS1 = L(U0) + sources, plus the per-direction max wave
//           speed (block reduce + one 64-bit atomic max per block).
// k_stage2: dt from the speeds, U1 = clamp(friction(U0 + dt*RHS1)) rebuilt on
//           tile+halo (never written to HBM; ghost halo cells through the
//           boundary-condition source map), then RHS2 = L(U1) + sources and
//           U2 = clamp(friction(0.5*(U0 + (U1 + dt*RHS2)), U1)) on the tile.
// The blocks at the domain corners also finish the corner ghost cells; the
// other ghost cells of U2 are filled by the boundary-condition kernel.
// Used when every ghost cell's boundary-condition source is an interior cell
// (WallBC and PeriodicBC on grids at least gc cells wide) and gc >= 2.
#include "swe_kernels.hpp"

namespace sfv {

constexpr int BOX_X = FT_X + 4;                 // state box: tile + 2-cell halo
constexpr int BOX_Y = FT_Y + 4;
constexpr int BB_X = FT_X + 5;                  // bottom corners of the box
constexpr int BB_Y = FT_Y + 5;
constexpr int NFX = (FT_X + 1) * FT_Y;          // x faces of a tile
constexpr int NFY = FT_X * (FT_Y + 1);          // y faces of a tile
constexpr int NF = NFX > NFY ? NFX : NFY;
constexpr int CELLS = FT_X * FT_Y;
constexpr int CPT = CELLS / F_THREADS;          // cells per thread
static_assert(CELLS % F_THREADS == 0, "tile must be a multiple of the block");

template <class S>
struct FusedSmem {
    S u[3][BOX_Y][BOX_X];
    double b[BB_Y][BB_X];
    double f[5][NF];        // face flux (3), depth of the minus side, depth of the plus side
};

__device__ __forceinline__ unsigned long long wave_max_u64f(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

struct TileGeom {
    int i0, j0;     // first interior cell of the tile
    int tw, th;     // valid tile width / height (partial tiles at the domain edge)
};

__device__ __forceinline__ TileGeom tile_geom(const KParams& P) {
    TileGeom t;
    t.i0 = P.gc + blockIdx.x * FT_X;
    t.j0 = P.gc + blockIdx.y * FT_Y;
    t.tw = min(FT_X, P.nxT - P.gc - t.i0);
    t.th = min(FT_Y, P.nyT - P.gc - t.j0);
    return t;
}

// Load the bottom corners of the box: corner (i0-2+x, j0-2+y) for the
// rows/columns that exist.
__device__ __forceinline__ void load_B(const KParams& P, const TileGeom& T, const double* __restrict__ B,
                                       double (*sb)[BB_X]) {
    const int nbx = P.nxT + 1, nby = P.nyT + 1;
    for (int e = threadIdx.x; e < BB_X * BB_Y; e += F_THREADS) {
        const int x = e % BB_X, y = e / BB_X;
        const int ci = T.i0 - 2 + x, cj = T.j0 - 2 + y;
        double v = 0.0;
        if (ci >= 0 && ci < nbx && cj >= 0 && cj < nby) v = B[(size_t)ci + (size_t)nbx * cj];
        sb[y][x] = v;
    }
}

// Is box cell (x, y) part of the cross-shaped stencil region of the tile?
__device__ __forceinline__ bool in_cross(const TileGeom& T, int x, int y) {
    const bool rows = y >= 2 && y < 2 + T.th;        // tile rows: full halo in x
    const bool cols = x >= 2 && x < 2 + T.tw;        // tile columns: full halo in y
    return (rows && x < T.tw + 4) || (cols && y < T.th + 4);
}

// Evaluate the faces of one direction into s.f and return the block-local
// max speed contribution of this thread (bit pattern).
template <class S>
__device__ __forceinline__ unsigned long long faces(const KParams& P, const TileGeom& T, FusedSmem<S>& s, int d) {
    unsigned long long smax = 0;
    const int fw = d == 0 ? T.tw + 1 : T.tw;         // faces per row
    const int fh = d == 0 ? T.th : T.th + 1;
    const int nf = fw * fh;
    const int stride_w = d == 0 ? FT_X + 1 : FT_X;   // face-buffer row pitch
    for (int e = threadIdx.x; e < nf; e += F_THREADS) {
        const int fx = e % fw, fy = e / fw;
        // face between cells (a) and (b) along d; box coordinates of cell b
        const int bx = d == 0 ? fx + 2 : fx + 2;
        const int by = d == 0 ? fy + 2 : fy + 2;
        const int sx = d == 0 ? 1 : 0, sy = d == 0 ? 0 : 1;
        // cells b-2, b-1 (= a), b, b+1 along d
        S c[4][3];
#pragma unroll
        for (int o = 0; o < 4; ++o)
#pragma unroll
            for (int v = 0; v < 3; ++v) c[o][v] = s.u[v][by + (o - 2) * sy][bx + (o - 2) * sx];
        // bottom faces of cells a and b: box corner (x, y) is corner (i0-2+x, j0-2+y)
        const int ax = bx - sx, ay = by - sy;
        double BlA, BrA, BlB, BrB;
        if (d == 0) {
            BlA = 0.5 * (s.b[ay][ax] + s.b[ay + 1][ax]);
            BrA = 0.5 * (s.b[ay][ax + 1] + s.b[ay + 1][ax + 1]);
            BlB = 0.5 * (s.b[by][bx] + s.b[by + 1][bx]);
            BrB = 0.5 * (s.b[by][bx + 1] + s.b[by + 1][bx + 1]);
        } else {
            BlA = 0.5 * (s.b[ay][ax] + s.b[ay][ax + 1]);
            BrA = 0.5 * (s.b[ay + 1][ax] + s.b[ay + 1][ax + 1]);
            BlB = 0.5 * (s.b[by][bx] + s.b[by][bx + 1]);
            BrB = 0.5 * (s.b[by + 1][bx] + s.b[by + 1][bx + 1]);
        }
        S La[3], Ra[3], Lb[3], Rb[3];
        if (P.rec == 1) {
            reconstruct_cell(c[0], c[1], c[2], P.theta, BlA, BrA, La, Ra);
            reconstruct_cell(c[1], c[2], c[3], P.theta, BlB, BrB, Lb, Rb);
        } else {
            const double bca = bcell4(P, s.b[ay][ax], s.b[ay][ax + 1], s.b[ay + 1][ax], s.b[ay + 1][ax + 1]);
            const double bcb = bcell4(P, s.b[by][bx], s.b[by][bx + 1], s.b[by + 1][bx], s.b[by + 1][bx + 1]);
#pragma unroll
            for (int v = 0; v < 3; ++v) { Ra[v] = c[1][v]; Lb[v] = c[2][v]; }
            Ra[0] = (S)(c[1][0] - bca);
            Lb[0] = (S)(c[2][0] - bcb);
        }
        double F[3];
        const double sp = central_upwind(d, Ra, Lb, P.phys, F);
        const int slot = fy * stride_w + fx;
        s.f[0][slot] = F[0];
        s.f[1][slot] = F[1];
        s.f[2][slot] = F[2];
        s.f[3][slot] = (double)Ra[0];
        s.f[4][slot] = (double)Lb[0];
        const unsigned long long bits = (unsigned long long)__double_as_longlong((double)(S)sp);
        smax = bits > smax ? bits : smax;
    }
    return smax;
}

// Accumulate direction d's flux difference and bottom source into acc for
// the thread's cells (compute_flux!, numericalflux.jl:27; source_term_bottom.jl:38).
template <class S>
__device__ __forceinline__ void apply_faces(const KParams& P, const TileGeom& T, FusedSmem<S>& s, int d,
                                            S acc[CPT][3]) {
    const int stride_w = d == 0 ? FT_X + 1 : FT_X;
    const double inv = d == 0 ? P.inv_dx : P.inv_dy;
    const double dd = d == 0 ? P.dx : P.dy;
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = threadIdx.x + q * F_THREADS;
        const int cx = c % FT_X, cy = c / FT_X;
        if (cx >= T.tw || cy >= T.th) continue;
        const int fm = cy * stride_w + cx;                                  // minus face
        const int fp = d == 0 ? fm + 1 : fm + stride_w;                     // plus face
#pragma unroll
        for (int v = 0; v < 3; ++v) acc[q][v] = (S)(acc[q][v] - inv * (s.f[v][fp] - s.f[v][fm]));
        if (P.bottom) {
            const int bx = cx + 2, by = cy + 2;
            double Bl, Br;
            if (d == 0) {
                Bl = 0.5 * (s.b[by][bx] + s.b[by + 1][bx]);
                Br = 0.5 * (s.b[by][bx + 1] + s.b[by + 1][bx + 1]);
            } else {
                Bl = 0.5 * (s.b[by][bx] + s.b[by][bx + 1]);
                Br = 0.5 * (s.b[by + 1][bx] + s.b[by + 1][bx + 1]);
            }
            const S hR = (S)s.f[3][fp];   // right face depth of this cell
            const S hL = (S)s.f[4][fm];   // left face depth of this cell
            const S hsum = hR + hL;
            acc[q][1 + d] = (S)(acc[q][1 + d] + -P.phys.g * slope_div(Br - Bl, dd, d == 0 ? P.inv_dx : P.inv_dy, d == 0 ? P.dx_pow2 : P.dy_pow2) * (hsum / 2.0));
        }
    }
}

template <class S>
__device__ __forceinline__ void block_max2(unsigned long long sx, unsigned long long sy, StepSlot* slot) {
    __shared__ unsigned long long red[2][F_THREADS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    sx = wave_max_u64f(sx);
    sy = wave_max_u64f(sy);
    if (lane == 0) { red[0][wid] = sx; red[1][wid] = sy; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long mx = red[0][0], my = red[1][0];
        for (int w = 1; w < F_THREADS / 64; ++w) {
            mx = red[0][w] > mx ? red[0][w] : mx;
            my = red[1][w] > my ? red[1][w] : my;
        }
        if (mx) atomicMax(&slot->wsmax[0], mx);
        if (my) atomicMax(&slot->wsmax[1], my);
    }
}

// ------------------------------------------------------------------- stage 1
template <class S>
__global__ void __launch_bounds__(F_THREADS) k_stage1(KParams P, StepCtl* ctl, int slot, const S* __restrict__ U0,
                                                      S* __restrict__ RHS, const double* __restrict__ B) {
    if (!ctl->slot[slot].active) return;
    __shared__ FusedSmem<S> s;
    const TileGeom T = tile_geom(P);
    const size_t n = (size_t)P.nxT * P.nyT;
    load_B(P, T, B, s.b);
    // state box.  A ghost cell is read through the boundary-condition maps,
    // as its source cell's value with the normal momentum negated at a wall:
    // the value update_bc! would have stored (the tiled kernels run only when
    // the sources are interior cells), so the state's ghost cells in memory
    // are never read.  The cross-shaped box holds no domain corner.
    const int gc = P.gc, nxT = P.nxT, nyT = P.nyT;
    for (int e = threadIdx.x; e < BOX_X * BOX_Y; e += F_THREADS) {
        const int x = e % BOX_X, y = e / BOX_X;
        if (!in_cross(T, x, y)) continue;
        int i = T.i0 - 2 + x, j = T.j0 - 2 + y;
        int neg = -1;
        if (i < gc) { neg = P.flip_lo_x[i] ? 1 : -1; i = P.src_lo_x[i]; }
        else if (i >= nxT - gc) { neg = P.flip_hi_x[i - (nxT - gc)] ? 1 : -1; i = P.src_hi_x[i - (nxT - gc)]; }
        else if (j < gc) { neg = P.flip_lo_y[j] ? 2 : -1; j = P.src_lo_y[j]; }
        else if (j >= nyT - gc) { neg = P.flip_hi_y[j - (nyT - gc)] ? 2 : -1; j = P.src_hi_y[j - (nyT - gc)]; }
        const size_t k = (size_t)i + (size_t)nxT * j;
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const S u = U0[v * n + k];
            s.u[v][y][x] = v == neg ? (S)(-u) : u;
        }
    }
    __syncthreads();
    S acc[CPT][3];
#pragma unroll
    for (int q = 0; q < CPT; ++q) acc[q][0] = acc[q][1] = acc[q][2] = S(0);
    const unsigned long long sx = faces(P, T, s, 0);
    __syncthreads();
    apply_faces(P, T, s, 0, acc);
    __syncthreads();
    const unsigned long long sy = faces(P, T, s, 1);
    __syncthreads();
    apply_faces(P, T, s, 1, acc);
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = threadIdx.x + q * F_THREADS;
        const int cx = c % FT_X, cy = c / FT_X;
        if (cx >= T.tw || cy >= T.th) continue;
        const size_t k = (size_t)(T.i0 + cx) + (size_t)P.nxT * (T.j0 + cy);
        acc[q][0] = apply_sources(P, ctl->slot[slot], acc[q][0], k);
#pragma unroll
        for (int v = 0; v < 3; ++v) RHS[v * n + k] = acc[q][v];
    }
    block_max2<S>(sx, sy, &ctl->slot[slot]);
}

// Stage-1 finish of one cell (rungekutta.jl:38-41 with dt, friction.jl:80-84
// from U0, timestepper.jl:53-63), from global memory.
template <class S>
__device__ __forceinline__ void finish1(const KParams& P, S x0, S x1, S x2, S r0, S r1, S r2, double Bm,
                                        double dt_d, S dt_s, S o[3]) {
    o[0] = (S)(x0 + dt_d * r0);
    o[1] = (S)(x1 + dt_d * r1);
    o[2] = (S)(x2 + dt_d * r2);
    if (P.friction) {
        const double fs = friction_scalar(P.friction, P.Cz, P.phys.kappa, x0, x1, x2, Bm);
        apply_friction(o, dt_s, fs);
    }
    if (o[0] - Bm < P.phys.cutoff) {
        o[0] = (S)jl_max((double)o[0], Bm);
        o[1] = (S)0.0;
        o[2] = (S)0.0;
    }
}

// ------------------------------------------------------------------- stage 2
template <class S>
__global__ void __launch_bounds__(F_THREADS) k_stage2(KParams P, StepCtl* ctl, int slot, const S* __restrict__ U0,
                                                      const S* __restrict__ RHS1, S* __restrict__ U2,
                                                      const double* __restrict__ B, int close) {
    const StepSlot& sl = ctl->slot[slot];
    const bool closer = close && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0;
    if (!sl.active) {
        // an inactive step passes inactivity on, so steps launched past the
        // end of a run stay no-ops
        if (closer) ctl->slot[1 - slot].active = 0;
        return;
    }
    __shared__ FusedSmem<S> s;
    const TileGeom T = tile_geom(P);
    const size_t n = (size_t)P.nxT * P.nyT;
    const double dt_d = stage1_dt<S>(P, ctl, sl);
    const S dt_s = (S)dt_d;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        ctl->dt = (double)dt_s;
        ctl->dt_stage1 = dt_d;
        // the step closes here (t += dt, next slot armed): the other threads
        // read only this step's slot, the next launches the next one
        if (close) close_step<S>(P, ctl, slot);
    }
    load_B(P, T, B, s.b);
    // U1 on the cross-shaped box.  A ghost cell takes the post-BC value: the
    // pre-BC value of its source cell (an interior cell for wall and
    // periodic boundaries) with the normal momentum negated at a wall.
    const int nxT = P.nxT, nyT = P.nyT, gc = P.gc;
    for (int e = threadIdx.x; e < BOX_X * BOX_Y; e += F_THREADS) {
        const int x = e % BOX_X, y = e / BOX_X;
        if (!in_cross(T, x, y)) continue;
        int i = T.i0 - 2 + x, j = T.j0 - 2 + y;
        int neg = -1;
        if (i < gc) { neg = P.flip_lo_x[i] ? 1 : -1; i = P.src_lo_x[i]; }
        else if (i >= nxT - gc) { neg = P.flip_hi_x[i - (nxT - gc)] ? 1 : -1; i = P.src_hi_x[i - (nxT - gc)]; }
        else if (j < gc) { neg = P.flip_lo_y[j] ? 2 : -1; j = P.src_lo_y[j]; }
        else if (j >= nyT - gc) { neg = P.flip_hi_y[j - (nyT - gc)] ? 2 : -1; j = P.src_hi_y[j - (nyT - gc)]; }
        const size_t k = (size_t)i + (size_t)nxT * j;
        const double Bm = B_cell(P, B, i, j);
        S o[3];
        finish1<S>(P, U0[k], U0[n + k], U0[2 * n + k], RHS1[k], RHS1[n + k], RHS1[2 * n + k], Bm, dt_d, dt_s, o);
        if (neg >= 0) o[neg] = (S)(-o[neg]);
        s.u[0][y][x] = o[0];
        s.u[1][y][x] = o[1];
        s.u[2][y][x] = o[2];
    }
    __syncthreads();
    S acc[CPT][3];
#pragma unroll
    for (int q = 0; q < CPT; ++q) acc[q][0] = acc[q][1] = acc[q][2] = S(0);
    faces(P, T, s, 0);
    __syncthreads();
    apply_faces(P, T, s, 0, acc);
    __syncthreads();
    faces(P, T, s, 1);
    __syncthreads();
    apply_faces(P, T, s, 1, acc);
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = threadIdx.x + q * F_THREADS;
        const int cx = c % FT_X, cy = c / FT_X;
        if (cx >= T.tw || cy >= T.th) continue;
        const int i = T.i0 + cx, j = T.j0 + cy;
        const size_t k = (size_t)i + (size_t)nxT * j;
        acc[q][0] = apply_sources(P, ctl->slot[slot], acc[q][0], k);
        const S x0 = s.u[0][cy + 2][cx + 2], x1 = s.u[1][cy + 2][cx + 2], x2 = s.u[2][cy + 2][cx + 2];
        S o[3];
        o[0] = (S)(x0 + dt_s * acc[q][0]);
        o[1] = (S)(x1 + dt_s * acc[q][1]);
        o[2] = (S)(x2 + dt_s * acc[q][2]);
        o[0] = (S)(0.5 * (U0[k] + o[0]));
        o[1] = (S)(0.5 * (U0[n + k] + o[1]));
        o[2] = (S)(0.5 * (U0[2 * n + k] + o[2]));
        const int bx = cx + 2, by = cy + 2;
        const double Bm = bcell4(P, s.b[by][bx], s.b[by][bx + 1], s.b[by + 1][bx], s.b[by + 1][bx + 1]);
        if (P.friction) {
            const double fs = friction_scalar(P.friction, P.Cz, P.phys.kappa, x0, x1, x2, Bm);
            apply_friction(o, dt_s, fs);
        }
        if (o[0] - Bm < P.phys.cutoff) {
            o[0] = (S)jl_max((double)o[0], Bm);
            o[1] = (S)0.0;
            o[2] = (S)0.0;
        }
        U2[k] = o[0];
        U2[n + k] = o[1];
        U2[2 * n + k] = o[2];
    }
    // The gc x gc corner cells of the domain are never written by the
    // boundary condition (bc.jl:92-109 skips them) and never read by a
    // stencil, but they are part of the state: both stages run the sources,
    // RK update, friction and clamp on them (rungekutta.jl:38-49).  The block
    // at each domain corner finishes that corner.
    const bool left = blockIdx.x == 0, right = blockIdx.x == gridDim.x - 1;
    const bool bottom = blockIdx.y == 0, top = blockIdx.y == gridDim.y - 1;
    if ((left || right) && (bottom || top)) {
        for (int e = threadIdx.x; e < 4 * gc * gc; e += F_THREADS) {
            const int corner = e / (gc * gc), r = e % (gc * gc);
            const bool hi_x = corner & 1, hi_y = corner >> 1;
            if ((hi_x ? !right : !left) || (hi_y ? !top : !bottom)) continue;
            const int i = hi_x ? nxT - gc + r % gc : r % gc;
            const int j = hi_y ? nyT - gc + r / gc : r / gc;
            const size_t k = (size_t)i + (size_t)nxT * j;
            const double Bm = B_cell(P, B, i, j);
            const S src = apply_sources(P, ctl->slot[slot], S(0), k);
            const S x0 = U0[k], x1 = U0[n + k], x2 = U0[2 * n + k];
            S u1[3];
            finish1<S>(P, x0, x1, x2, src, S(0), S(0), Bm, dt_d, dt_s, u1);
            S o[3];
            o[0] = (S)(u1[0] + dt_s * src);
            o[1] = (S)(u1[1] + dt_s * S(0));
            o[2] = (S)(u1[2] + dt_s * S(0));
            o[0] = (S)(0.5 * (x0 + o[0]));
            o[1] = (S)(0.5 * (x1 + o[1]));
            o[2] = (S)(0.5 * (x2 + o[2]));
            if (P.friction) {
                const double fs = friction_scalar(P.friction, P.Cz, P.phys.kappa, u1[0], u1[1], u1[2], Bm);
                apply_friction(o, dt_s, fs);
            }
            if (o[0] - Bm < P.phys.cutoff) {
                o[0] = (S)jl_max((double)o[0], Bm);
                o[1] = (S)0.0;
                o[2] = (S)0.0;
            }
            U2[k] = o[0];
            U2[n + k] = o[1];
            U2[2 * n + k] = o[2];
        }
    }
}

template <class S>
hipError_t launch_stage1(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, const S* U0, S* RHS,
                         const double* B) {
    const dim3 grid((P.nx + FT_X - 1) / FT_X, (P.ny + FT_Y - 1) / FT_Y);
    hipLaunchKernelGGL(k_stage1<S>, grid, dim3(F_THREADS), 0, st, P, ctl, slot, U0, RHS, B);
    return hipPeekAtLastError();
}
template <class S>
hipError_t launch_stage2(hipStream_t st, const KParams& P, StepCtl* ctl, int slot, const S* U0, const S* RHS1, S* U2,
                         const double* B, int close) {
    const dim3 grid((P.nx + FT_X - 1) / FT_X, (P.ny + FT_Y - 1) / FT_Y);
    hipLaunchKernelGGL(k_stage2<S>, grid, dim3(F_THREADS), 0, st, P, ctl, slot, U0, RHS1, U2, B, close);
    return hipPeekAtLastError();
}
template hipError_t launch_stage1<double>(hipStream_t, const KParams&, StepCtl*, int, const double*, double*,
                                          const double*);
template hipError_t launch_stage1<float>(hipStream_t, const KParams&, StepCtl*, int, const float*, float*,
                                         const double*);
template hipError_t launch_stage2<double>(hipStream_t, const KParams&, StepCtl*, int, const double*, const double*,
                                          double*, const double*, int);
template hipError_t launch_stage2<float>(hipStream_t, const KParams&, StepCtl*, int, const float*, const float*,
                                         float*, const double*, int);

}  // namespace sfv
