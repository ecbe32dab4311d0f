// swe_adjoint.hip — reverse-mode gradient of a time-stepped shallow-water run.
//
// J = sum over interior cells of (h_T - h_target)^2, h = w - B_cell, after
// nsteps RK2 steps (perform_step!(sim, Inf) each, simulator.jl:145-177).
// Gradients: dJ/dB (every bottom corner), dJ/dCz (ImplicitFriction, Cz =
// Manning n^2), dJ/dU0 (interior initial state).  The derivative is the one
// forward-mode differentiation of the reference gives: every branch the
// forward step took (wet/dry cut-offs, minmod, desingularisation, the dry
// clamp, the argmax face of the CFL time step) is taken again and
// differentiated; dt is differentiated through its maximum wave speed.
//
// Memory: the forward pass keeps a checkpoint every `checkpoint_every`
// steps.  The reverse pass recomputes one segment from its checkpoint (the
// per-pass kernels, bitwise the forward run), keeping each step's stage-1
// state U1, right-hand sides R1 = L(U0) + sources, R2 = L(U1) + sources and
// control block when they fit (on a second stream, overlapped with the
// previous segment's reverse steps, when two sets fit), then walks it
// backwards.  Per stage in reverse order a reverse step applies the
// boundary-condition fold (read as a gather), dry clamp, implicit friction
// and the RK combination, then the face fluxes of both directions (the
// central-upwind flux swept back by hand at each face) and the
// reconstruction (its branch-wise linear Jacobian).  All kernels are gathers
// apart from the dt-face search: every output element is written by one
// thread.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sinfvm_hip.h"
#include "sfv_handle.hpp"
#include "swe_dual.hpp"

namespace sfv {
namespace ad {

// ---- the step's arithmetic on a generic number type T (double or Dual<N>),
// value-for-value the same operations as swe_physics.hpp.

template <class T>
__device__ T desingularize(const T& h, double kappa) {
    const double ah = __builtin_fabs(val(h));
    if (!(ah < kappa)) return h;
    const T h2 = h * h;
    const double sgn = __builtin_signbit(val(h)) ? -1.0 : 1.0;
    const T inner = h2 / (2 * kappa) + kappa / 2.0;
    const T mn = djl_min(inner, T(kappa));
    const T mx = djl_max(dabs(h), mn);
    return sgn * mx;
}

template <class T>
__device__ void phys(int d, const T& h, const T& hu, const T& hv, const PhysParams& p, T f[3], T e[2]) {
    const T hs = ad::desingularize(h, p.kappa);
    const T u = hu / hs;
    const T v = hv / hs;
    const T h2 = h * h;
    const T c = dsqrt(p.g * h);
    if (d == 0) {
        f[0] = p.rho * h * u;
        f[1] = p.rho * h * u * u + 0.5 * p.rho * p.g * h2;
        f[2] = p.rho * h * u * v;
        e[0] = u + c;
        e[1] = u - c;
    } else {
        f[0] = p.rho * h * v;
        f[1] = p.rho * h * u * v;
        f[2] = p.rho * h * v * v + 0.5 * p.rho * p.g * h2;
        e[0] = v + c;
        e[1] = v - c;
    }
}

// centralupwind.jl:48-77; returns the face speed, F the flux.
template <class T>
__device__ T central_upwind(int d, const T m[3], const T q[3], const PhysParams& p, T F[3]) {
    T fm[3] = {T(0.0), T(0.0), T(0.0)}, em[2] = {T(0.0), T(0.0)};
    T fp[3] = {T(0.0), T(0.0), T(0.0)}, ep[2] = {T(0.0), T(0.0)};
    if (val(m[0]) > p.cutoff) ad::phys(d, m[0], m[1], m[2], p, fm, em);
    if (val(q[0]) > p.cutoff) ad::phys(d, q[0], q[1], q[2], p, fp, ep);
    const T aplus = djl_max(djl_max(ep[0], em[0]), T(0.0));
    const T aminus = djl_min(djl_min(ep[1], em[1]), T(0.0));
    const T da = aplus - aminus;
    if (__builtin_fabs(val(da)) < p.kappa) {
        F[0] = F[1] = F[2] = T(0.0);
        return T(0.0);
    }
    const T pm = (aplus * aminus) / da;
    for (int v = 0; v < 3; ++v) {
        const T diff = q[v] - m[v];
        F[v] = (aplus * fm[v] - aminus * fp[v]) / da + pm * diff;
    }
    if (val(m[0]) < p.cutoff && val(q[0]) < p.cutoff) return T(0.0);
    return djl_max(dabs(aplus), dabs(aminus));
}

// Which operand Julia's max / min returns (0: x, 1: y), as djl_max / djl_min.
__device__ __forceinline__ int jmax_pick(double x, double y) {
    if (x != x) return 0;
    if (y != y) return 1;
    if (x > y) return 0;
    if (y > x) return 1;
    return __builtin_signbit(x) ? 1 : 0;
}
__device__ __forceinline__ int jmin_pick(double x, double y) {
    if (x != x) return 0;
    if (y != y) return 1;
    if (x < y) return 0;
    if (y < x) return 1;
    return __builtin_signbit(x) ? 0 : 1;
}

// d desingularize(h)/dh along the branch the value took (equation.jl:40-67;
// the operations and picks of ad::desingularize).
__device__ __forceinline__ double desing_dh(double h, double kappa) {
    if (!(__builtin_fabs(h) < kappa)) return 1.0;
    const double inner = h * h / (2 * kappa) + kappa / 2.0;
    const int mn_k = jmin_pick(inner, kappa);                  // 1: min returned kappa
    const double mn = mn_k ? kappa : inner;
    const double dmn = mn_k ? 0.0 : (h + h) * (1.0 / (2 * kappa));
    const double sg = __builtin_signbit(h) ? -1.0 : 1.0;
    const int mx_mn = jmax_pick(__builtin_fabs(h), mn);      // 1: max returned the min
    return sg * (mx_mn ? dmn : sg);
}

// One side of a face: its flux f and eigenvalues e (swe_2D.jl:43-81), zero
// for a dry side (centralupwind.jl:51-61 leaves them at zero).
__device__ __forceinline__ bool side_values(int d, const double s[3], const PhysParams& p, double f[3], double e[2]) {
    for (int v = 0; v < 3; ++v) f[v] = 0.0;
    e[0] = e[1] = 0.0;
    if (!(s[0] > p.cutoff)) return false;
    ad::phys<double>(d, s[0], s[1], s[2], p, f, e);
    return true;
}

// Reverse sweep of one wet side: given the adjoints lf of its flux and le of
// its eigenvalues, the adjoint of its state (h, hu, hv).  With the normal
// momentum n = 1 + d and the tangential t = 2 - d, the side computes
// un = s_n / hs, ut = s_t / hs, hs = desingularize(h), c = sqrt(g h),
// f = [rho h un, rho h un^2 + 0.5 rho g h^2 (at n), rho h un ut (at t)],
// e = [un + c, un - c].
__device__ __forceinline__ void side_vjp(int d, const double s[3], const PhysParams& p, const double lf[3],
                                         const double le[2], double ls[3]) {
    const double h = s[0];
    const double sn = d == 0 ? s[1] : s[2], st = d == 0 ? s[2] : s[1];
    const double ln = d == 0 ? lf[1] : lf[2], lt = d == 0 ? lf[2] : lf[1];
    const double hs = ad::desingularize<double>(h, p.kappa);
    const double un = sn / hs, ut = st / hs;
    const double c = sqrt(p.g * h);
    const double rh = p.rho * h;
    const double l_un = lf[0] * rh + ln * (2.0 * rh * un) + lt * (rh * ut) + le[0] + le[1];
    const double l_ut = lt * (rh * un);
    double l_h = lf[0] * (p.rho * un) + ln * (p.rho * un * un + p.rho * p.g * h) + lt * (p.rho * un * ut);
    if (c > 0.0) l_h += (le[0] - le[1]) * (p.g * (0.5 / c));
    const double ihs = 1.0 / hs;
    const double l_hs = -(l_un * un + l_ut * ut) * ihs;
    ls[0] = l_h + l_hs * desing_dh(h, p.kappa);
    const double l_sn = l_un * ihs, l_st = l_ut * ihs;
    ls[1] = d == 0 ? l_sn : l_st;
    ls[2] = d == 0 ? l_st : l_sn;
}

// Reverse sweep of the central-upwind flux (centralupwind.jl:48-77) at one
// face: given the flux adjoint lF (and the adjoint l_speed of the face's
// speed), the adjoints of the minus side m and the plus side q, and the
// face's speed.  The values and every branch are those of central_upwind on
// doubles; the combination of the two sides and each side's flux are
// differentiated by hand (side_vjp).
__device__ __forceinline__ void face_vjp(int d, const double m[3], const double q[3], const PhysParams& p, const double lF[3],
                         double lm[3], double lq[3], double& speed, double l_speed = 0.0) {
    double fm[3], em[2], fp[3], ep[2];
    const bool wm = side_values(d, m, p, fm, em);
    const bool wq = side_values(d, q, p, fp, ep);
    for (int v = 0; v < 3; ++v) lm[v] = lq[v] = 0.0;
    const int pk1 = jmax_pick(ep[0], em[0]);
    const double ap1 = pk1 ? em[0] : ep[0];
    const int pk2 = jmax_pick(ap1, 0.0);
    const double aplus = pk2 ? 0.0 : ap1;
    const int mk1 = jmin_pick(ep[1], em[1]);
    const double am1 = mk1 ? em[1] : ep[1];
    const int mk2 = jmin_pick(am1, 0.0);
    const double aminus = mk2 ? 0.0 : am1;
    const double D = aplus - aminus;
    // the face speed central_upwind returns (centralupwind.jl:48-77)
    speed = (__builtin_fabs(D) < p.kappa || (m[0] < p.cutoff && q[0] < p.cutoff))
                ? 0.0 : djl_max(__builtin_fabs(aplus), __builtin_fabs(aminus));
    if (__builtin_fabs(D) < p.kappa) return;       // the flux is the constant zero
    const double pm = (aplus * aminus) / D;
    double l_ap = 0.0, l_am = 0.0, l_D = 0.0, l_pm = 0.0, l_fm[3], l_fp[3];
    // the speed's own adjoint: speed = max(|a+|, |a-|) unless both sides are dry
    if (l_speed != 0.0 && !(m[0] < p.cutoff && q[0] < p.cutoff)) {
        if (jmax_pick(__builtin_fabs(aplus), __builtin_fabs(aminus)))
            l_am += l_speed * (__builtin_signbit(aminus) ? -1.0 : 1.0);
        else
            l_ap += l_speed * (__builtin_signbit(aplus) ? -1.0 : 1.0);
    }
    for (int v = 0; v < 3; ++v) {
        const double N = aplus * fm[v] - aminus * fp[v];
        const double lN = lF[v] / D;
        l_D -= lN * (N / D);
        l_pm += lF[v] * (q[v] - m[v]);
        lq[v] += lF[v] * pm;
        lm[v] -= lF[v] * pm;
        l_ap += lN * fm[v];
        l_am -= lN * fp[v];
        l_fm[v] = lN * aplus;
        l_fp[v] = -lN * aminus;
    }
    l_ap += l_pm * aminus / D;
    l_am += l_pm * aplus / D;
    l_D -= l_pm * pm / D;
    l_ap += l_D;
    l_am -= l_D;
    double l_ep[2] = {0.0, 0.0}, l_em[2] = {0.0, 0.0};
    // explicit branches (a conditional lvalue would put the arrays in memory)
    if (!pk2) {
        if (pk1) l_em[0] += l_ap;
        else l_ep[0] += l_ap;
    }
    if (!mk2) {
        if (mk1) l_em[1] += l_am;
        else l_ep[1] += l_am;
    }
    double gm[3] = {0.0, 0.0, 0.0}, gq[3] = {0.0, 0.0, 0.0};
    if (wm) side_vjp(d, m, p, l_fm, l_em, gm);
    if (wq) side_vjp(d, q, p, l_fp, l_ep, gq);
    for (int a = 0; a < 3; ++a) {
        lm[a] += gm[a];
        lq[a] += gq[a];
    }
}

// linearreconstruction.jl:28-35: the argument minmod selects (0 if none), ties
// resolved as Julia's min/max resolve them.
template <class T>
__device__ T minmod(const T& a, const T& b, const T& c) {
    const double x = val(a), y = val(b), z = val(c);
    if (x != x || y != y || z != z) return T(0.0);
    if (x > 0 && y > 0 && z > 0) return djl_min(djl_min(a, b), c);
    if (x < 0 && y < 0 && z < 0) return djl_max(djl_max(a, b), c);
    return T(0.0);
}

// minmod_slope + dry fix + face states (linearreconstruction.jl:37-43, 73-100).
template <class T>
__device__ void reconstruct(const T l[3], const T c[3], const T r[3], double theta, const T& Bl, const T& Br, T L[3],
                            T R[3]) {
    T slope[3];
    for (int v = 0; v < 3; ++v) {
        const T fwd = r[v] - c[v];
        const T bwd = c[v] - l[v];
        const T central = (fwd + bwd) / 2.0;
        slope[v] = ad::minmod(theta * fwd, central, theta * bwd);
    }
    const T w = c[0];
    if (val(w - 0.5 * slope[0]) < val(Bl)) slope[0] = 2.0 * (w - Bl);
    else if (val(w + 0.5 * slope[0]) < val(Br)) slope[0] = 2.0 * (Br - w);
    for (int v = 0; v < 3; ++v) {
        L[v] = c[v] - 0.5 * slope[v];
        R[v] = c[v] + 0.5 * slope[v];
    }
    L[0] = L[0] - Bl;
    R[0] = R[0] - Br;
}

// friction.jl:30-43, 62-78: the implicit friction scalar from the stage input.
template <class T>
__device__ T friction_scalar(int law, const T& Cz, double kappa, const T& w, const T& hu, const T& hv, const T& Bm) {
    const T hs = ad::desingularize(w - Bm, kappa);
    const T u = hu / hs, v = hv / hs;
    T speed = dsqrt(u * u + v * v);
    if (val(speed) == 0.0) speed = T(0.0);
    if (law == 2) { const T denom = dcbrt(hs) * hs; return -Cz * speed / denom; }
    if (law == 3) { const T denom = dcbrt(hs) * hs * hs; return -Cz * speed / denom; }
    const T denom = hs * hs;
    return -Cz * speed / denom;
}

__device__ inline double Bc4(const KParams& P, const double* B, int i, int j) {
    return 0.25 * (((Bcorner(P, B, i, j) + Bcorner(P, B, i + 1, j)) + Bcorner(P, B, i, j + 1)) + Bcorner(P, B, i + 1, j + 1));
}

// Cell (i,j)'s face states in direction d (linearreconstruction.jl:73-100).
template <class T>
__device__ void recon_at(const KParams& P, const double* U, const double* B, int d, int i, int j, T L[3], T R[3]) {
    const size_t n = (size_t)P.nxT * P.nyT;
    const int di = d == 0 ? 1 : 0, dj = d == 0 ? 0 : 1;
    T cl[3], cc[3], cr[3];
    for (int v = 0; v < 3; ++v) {
        cl[v] = T(U[v * n + cid(P, i - di, j - dj)]);
        cc[v] = T(U[v * n + cid(P, i, j)]);
        cr[v] = T(U[v * n + cid(P, i + di, j + dj)]);
    }
    const T Bl(B_face_left(P, B, i, j, d)), Br(B_face_right(P, B, i, j, d));
    ad::reconstruct(cl, cc, cr, P.theta, Bl, Br, L, R);
}

// Face states of cell (i,j) along d, as recon_at<double> gives them, and the
// branch each slope took (linearreconstruction.jl:28-43, 81-91): pick[v] is
// 0 for theta*fwd, 1 for (fwd+bwd)/2, 2 for theta*bwd, 3 for minmod's zero;
// dry is 0, or 1 for the 2(w - Bl) line, 2 for the 2(Br - w) line.  Ties
// resolve as Julia's min/max (and djl_min/djl_max) resolve them.  Every
// branch is linear in its stencil, and variable v's face states depend on
// variable v only, so the picks give the reconstruction's Jacobian.
__device__ void recon_pick(const KParams& P, const double* U, const double* B, int d, int i, int j, double L[3],
                           double R[3], int pick[3], int& dry) {
    const size_t n = (size_t)P.nxT * P.nyT;
    const int di = d == 0 ? 1 : 0, dj = d == 0 ? 0 : 1;
    const double theta = P.theta;
    double cc[3], slope[3];
    for (int v = 0; v < 3; ++v) {
        const double l = U[v * n + cid(P, i - di, j - dj)];
        const double c = U[v * n + cid(P, i, j)];
        const double r = U[v * n + cid(P, i + di, j + dj)];
        cc[v] = c;
        const double fwd = r - c, bwd = c - l;
        const double a = theta * fwd, b = (fwd + bwd) / 2.0, e = theta * bwd;
        int k = 3;
        double m = 0.0;
        if (a != a || b != b || e != e) {
        } else if (a > 0 && b > 0 && e > 0) {        // djl_min(djl_min(a, b), e)
            k = b <= a ? 1 : 0;
            m = k ? b : a;
            if (e <= m) { k = 2; m = e; }
        } else if (a < 0 && b < 0 && e < 0) {        // djl_max(djl_max(a, b), e)
            k = b >= a ? 1 : 0;
            m = k ? b : a;
            if (e >= m) { k = 2; m = e; }
        }
        pick[v] = k;
        slope[v] = m;
    }
    const double Bl = B_face_left(P, B, i, j, d), Br = B_face_right(P, B, i, j, d);
    const double w = cc[0];
    dry = 0;
    if (w - 0.5 * slope[0] < Bl) { slope[0] = 2.0 * (w - Bl); dry = 1; }
    else if (w + 0.5 * slope[0] < Br) { slope[0] = 2.0 * (Br - w); dry = 2; }
    for (int v = 0; v < 3; ++v) {
        L[v] = cc[v] - 0.5 * slope[v];
        R[v] = cc[v] + 0.5 * slope[v];
    }
    L[0] = L[0] - Bl;
    R[0] = R[0] - Br;
}

// d slope / d U at stencil position p (-1 left, 0 centre, 1 right) for the
// branch pick (and, for w, the dry-fix line).
__device__ __forceinline__ double slope_coef(int pick, int dry, int p, double theta) {
    if (dry == 1) return p == 0 ? 2.0 : 0.0;
    if (dry == 2) return p == 0 ? -2.0 : 0.0;
    if (pick == 0) return p == 1 ? theta : (p == 0 ? -theta : 0.0);
    if (pick == 1) return p == 1 ? 0.5 : (p == -1 ? -0.5 : 0.0);
    if (pick == 2) return p == 0 ? theta : (p == -1 ? -theta : 0.0);
    return 0.0;
}

// ---------------------------------------------------------------- kernels

// Boundary-condition fold (adjoint of update_bc!): ghost adjoints go to the
// cells the ghosts were copied from (sign flipped with the momentum), then
// ghosts are zeroed.  axis 0: x ghosts, one thread per interior row; axis 1:
// y ghosts, one thread per interior column.  The gc x gc corners are never
// read by the step and are zeroed too.
__global__ void k_adj_bc(KParams P, double* __restrict__ L, int axis) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)P.nxT * P.nyT;
    const int gc = P.gc;
    if (axis == 0) {
        if (t >= P.ny) return;
        const int j = gc + t;
        for (int g = 0; g < gc; ++g) {
            const size_t klo = cid(P, g, j), khi = cid(P, P.nxT - gc + g, j);
            const size_t slo = cid(P, P.src_lo_x[g], j), shi = cid(P, P.src_hi_x[g], j);
            for (int v = 0; v < 3; ++v) {
                const double flo = (P.flip_lo_x[g] && v == 1) ? -1.0 : 1.0;
                const double fhi = (P.flip_hi_x[g] && v == 1) ? -1.0 : 1.0;
                L[v * n + slo] += flo * L[v * n + klo];
                L[v * n + klo] = 0.0;
                L[v * n + shi] += fhi * L[v * n + khi];
                L[v * n + khi] = 0.0;
            }
        }
    } else {
        if (t >= P.nx) {
            // corners
            const int c = t - P.nx;
            if (c < 4 * gc * gc) {
                const int corner = c / (gc * gc), rr = c % (gc * gc);
                const int ci = (corner & 1) ? P.nxT - gc + rr % gc : rr % gc;
                const int cj = (corner >> 1) ? P.nyT - gc + rr / gc : rr / gc;
                for (int v = 0; v < 3; ++v) L[v * n + cid(P, ci, cj)] = 0.0;
            }
            return;
        }
        const int i = gc + t;
        for (int g = 0; g < gc; ++g) {
            const size_t klo = cid(P, i, g), khi = cid(P, i, P.nyT - gc + g);
            const size_t slo = cid(P, i, P.src_lo_y[g]), shi = cid(P, i, P.src_hi_y[g]);
            for (int v = 0; v < 3; ++v) {
                const double flo = (P.flip_lo_y[g] && v == 2) ? -1.0 : 1.0;
                const double fhi = (P.flip_hi_y[g] && v == 2) ? -1.0 : 1.0;
                L[v * n + slo] += flo * L[v * n + klo];
                L[v * n + klo] = 0.0;
                L[v * n + shi] += fhi * L[v * n + khi];
                L[v * n + khi] = 0.0;
            }
        }
    }
}

// The value k_adj_bc's two folds leave at cell (i,j), read without running
// them: ghosts and corners read 0; an interior cell reads its own adjoint
// plus, in the folds' order (x then y, by ghost index, low side then high
// side), the adjoints of the ghosts copied from it, sign-flipped with the
// normal momentum.  Exact because the ghosts' sources are interior cells
// (no ghost feeds another) and the x and y folds touch disjoint ghosts.
__device__ __forceinline__ void folded_adj(const KParams& P, const double* __restrict__ L, int i, int j, double lo[3]) {
    const size_t n = (size_t)P.nxT * P.nyT;
    const int gc = P.gc;
    lo[0] = lo[1] = lo[2] = 0.0;
    if (i < gc || i >= P.nxT - gc || j < gc || j >= P.nyT - gc) return;
    const size_t k = cid(P, i, j);
    for (int v = 0; v < 3; ++v) lo[v] = L[v * n + k];
    for (int g = 0; g < gc; ++g) {
        if (P.src_lo_x[g] == i) {
            const size_t q = cid(P, g, j);
            for (int v = 0; v < 3; ++v) lo[v] += ((P.flip_lo_x[g] && v == 1) ? -1.0 : 1.0) * L[v * n + q];
        }
        if (P.src_hi_x[g] == i) {
            const size_t q = cid(P, P.nxT - gc + g, j);
            for (int v = 0; v < 3; ++v) lo[v] += ((P.flip_hi_x[g] && v == 1) ? -1.0 : 1.0) * L[v * n + q];
        }
    }
    for (int g = 0; g < gc; ++g) {
        if (P.src_lo_y[g] == j) {
            const size_t q = cid(P, i, g);
            for (int v = 0; v < 3; ++v) lo[v] += ((P.flip_lo_y[g] && v == 2) ? -1.0 : 1.0) * L[v * n + q];
        }
        if (P.src_hi_y[g] == j) {
            const size_t q = cid(P, i, P.nyT - gc + g);
            for (int v = 0; v < 3; ++v) lo[v] += ((P.flip_hi_y[g] && v == 2) ? -1.0 : 1.0) * L[v * n + q];
        }
    }
}

// Block sum of one value into part[blockIdx]; deterministic (fixed tree).
__device__ void block_sum(double v, double* part, int slot) {
    __shared__ double red[256];
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    red[tid] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) part[slot] = red[0];
    __syncthreads();
}

// Adjoint of: pre = in + dt*R (stage 2: pre = 0.5*(U0 + in + dt*R)),
// friction (scalar from `in`), dry clamp.  Lout: adjoint of the stage
// output, read through its boundary-condition fold (folded_adj).  Stage 2 writes Lin and LU0 for every
// cell and resets the step's argmax search (*best); stage 1 adds to Lin.
// Writes LR = adjoint of R, adds the B_cell adjoint, and block partial sums
// of the dt and Cz adjoints.
__global__ void __launch_bounds__(256) k_adj_tail(KParams P, const StepCtl* ctl, int stage, const double* __restrict__ U0,
                                                  const double* __restrict__ Uin, const double* __restrict__ R,
                                                  const double* __restrict__ B, const double* __restrict__ Lout,
                                                  double* __restrict__ Lin, double* __restrict__ LU0,
                                                  double* __restrict__ LR, double* __restrict__ LBc,
                                                  double* __restrict__ part_dt, double* __restrict__ part_cz,
                                                  unsigned long long* __restrict__ best) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // the launchers' blocks are 64 wide, so a wave shares threadIdx.y
    const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * blockDim.y + threadIdx.y);
    const int nblk = gridDim.x * gridDim.y, bid = blockIdx.x + gridDim.x * blockIdx.y;
    if (stage == 2 && bid == 0 && threadIdx.x == 0 && threadIdx.y == 0) *best = ~0ull;
    double gdt = 0.0, gcz = 0.0;
    if (i < P.nxT && j < P.nyT) {
        const size_t n = (size_t)P.nxT * P.nyT;
        const size_t k = cid(P, i, j);
        const double dt = ctl->dt_stage1;
        double x[3], r[3], ax[3], pre[3], lo[3];
        for (int v = 0; v < 3; ++v) {
            x[v] = Uin[v * n + k];
            r[v] = R[v * n + k];
            ax[v] = x[v] + dt * r[v];
            pre[v] = stage == 2 ? 0.5 * (U0[v * n + k] + ax[v]) : ax[v];
        }
        folded_adj(P, Lout, i, j, lo);
        const double Bm = Bc4(P, B, i, j);
        double fs = 0.0, D = 1.0;
        if (P.friction) {
            fs = ad::friction_scalar<double>(P.friction, P.Cz, P.phys.kappa, x[0], x[1], x[2], Bm);
            D = 1 - dt * fs;
        }
        double o[3] = {pre[0] * (1 - dt * 0.0), pre[1] / D, pre[2] / D};
        double lBc = 0.0;
        // dry clamp (timestepper.jl:53-63)
        if (o[0] - Bm < P.phys.cutoff) {
            // jl_max(o0, Bm): o0 if larger, Bm if larger, on a tie Bm when o0 is -0.0
            const bool takes_w = o[0] > Bm || (!(Bm > o[0]) && !__builtin_signbit(o[0]));
            if (!takes_w) lBc += lo[0];
            lo[0] = takes_w ? lo[0] : 0.0;
            lo[1] = 0.0;
            lo[2] = 0.0;
        }
        // friction: o1 = pre1 / D, o2 = pre2 / D, D = 1 - dt*fs(x, Bm, Cz)
        double lpre[3] = {lo[0], lo[1] / D, lo[2] / D};
        // stage 2 is the first writer of Lin and LU0 in a reverse step: it
        // stores 0.0 + its terms, stage 1 adds to what is there
        double lin[3];
        for (int v = 0; v < 3; ++v) lin[v] = stage == 2 ? 0.0 : Lin[v * n + k];
        if (P.friction) {
            const double lD = -(lo[1] * pre[1] + lo[2] * pre[2]) / (D * D);
            const double lfs = lD * (-dt);
            gdt += lD * (-fs);
            using D5 = Dual<5>;
            const D5 fsd = ad::friction_scalar<D5>(P.friction, D5::var(P.Cz, 4), P.phys.kappa, D5::var(x[0], 0),
                                               D5::var(x[1], 1), D5::var(x[2], 2), D5::var(Bm, 3));
            for (int v = 0; v < 3; ++v) lin[v] += lfs * fsd.d[v];
            lBc += lfs * fsd.d[3];
            gcz += lfs * fsd.d[4];
        }
        // RK combination
        double lax[3];
        for (int v = 0; v < 3; ++v) {
            lax[v] = stage == 2 ? 0.5 * lpre[v] : lpre[v];
            if (stage == 2) LU0[v * n + k] = 0.0 + 0.5 * lpre[v];
            Lin[v * n + k] = lin[v] + lax[v];
            LR[v * n + k] = dt * lax[v];
            gdt += lax[v] * r[v];
        }
        LBc[k] += lBc;
    }
    block_sum(gdt, part_dt, (stage == 2 ? 0 : nblk) + bid);
    block_sum(gcz, part_cz, (stage == 2 ? 0 : nblk) + bid);
}

// The direction whose wave speed set dt: min(Δx/a_x, Δy/a_y) picks y on a
// tie (simulator.jl:147-150 with Julia's minimum).
__device__ __forceinline__ int dt_direction(const KParams& P, const StepCtl* ctl) {
    const double ax = __longlong_as_double((long long)ctl->slot[0].wsmax[0]);
    const double ay = __longlong_as_double((long long)ctl->slot[0].wsmax[1]);
    return (P.dx / ax < P.dy / ay) ? 0 : 1;
}

// Face fluxes of direction d (numericalflux.jl:21-33): face between cell c
// and c+1 along d, c from gc-1 to nT-gc-1, other index interior.  The flux
// adjoint is (1/Δ)(LR(c+1) - LR(c)) over interior cells; its transpose
// through the central-upwind Jacobian gives the adjoints of c's right face
// state and (c+1)'s left face state.
__global__ void __launch_bounds__(256) k_adj_faces(KParams P, int d, const double* __restrict__ U,
                                                   const double* __restrict__ B, const double* __restrict__ LR,
                                                   double* __restrict__ LfL, double* __restrict__ LfR,
                                                   const StepCtl* __restrict__ ctl, unsigned long long* __restrict__ best) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nT = d == 0 ? P.nxT : P.nyT, nO = d == 0 ? P.ny : P.nx;
    const int nf = nT - 2 * P.gc + 1;
    if (t >= nf * nO) return;
    // consecutive threads take consecutive i in both directions (coalesced)
    const int s = d == 0 ? P.gc - 1 + t % nf : P.gc - 1 + t / nO;
    const int o = d == 0 ? P.gc + t / nf : P.gc + t % nO;
    const int i = d == 0 ? s : o, j = d == 0 ? o : s;
    const int i1 = d == 0 ? i + 1 : i, j1 = d == 0 ? j : j + 1;
    const size_t n = (size_t)P.nxT * P.nyT;
    const size_t kc = cid(P, i, j), k1 = cid(P, i1, j1);
    const double inv = d == 0 ? P.inv_dx : P.inv_dy;
    const bool c_in = s >= P.gc, c1_in = s + 1 < nT - P.gc;
    double lF[3];
    for (int v = 0; v < 3; ++v) lF[v] = inv * ((c1_in ? LR[v * n + k1] : 0.0) - (c_in ? LR[v * n + kc] : 0.0));
    double Lc[3], Rc[3], L1[3], R1[3];
    recon_at<double>(P, U, B, d, i, j, Lc, Rc);
    recon_at<double>(P, U, B, d, i1, j1, L1, R1);
    double gm[3], gq[3], speed;
    ad::face_vjp(d, Rc, L1, P.phys, lF, gm, gq, speed);
    for (int a = 0; a < 3; ++a) {
        LfR[a * n + kc] = gm[a];
        LfL[a * n + k1] = gq[a];
    }
    // With best given: the face that set dt.  maximum(wavespeeds) returns the
    // first interior cell in memory order whose wavespeed, the max of its two
    // face speeds (numericalflux.jl:29-32), is the step's maximum; a cell's
    // wavespeed is the maximum exactly when one of its faces' speeds is, so
    // each face at the maximum offers its interior cells and the smallest
    // memory index wins.
    if (best && d == dt_direction(P, ctl) &&
        (unsigned long long)__double_as_longlong(speed) == ctl->slot[0].wsmax[d]) {
        unsigned long long c = ~0ull;
        if (c_in) c = kc;
        if (c1_in && k1 < c) c = k1;
        if (c != ~0ull) atomicMin(best, c);
    }
}

// Cell (i,j), direction d: adjoint of the reconstruction (and of the bottom
// source, source_term_bottom.jl:24-41) gathered from the face-state adjoints
// of the cell and its two neighbours along d.  Adds it to acc (has: whether
// the cell takes part in direction d); adds the face bottom adjoints of the
// cell to LBf (planes: x left, x right, y left, y right).
__device__ __forceinline__ bool recon_dir(const KParams& P, int d, int i, int j, const double* __restrict__ U,
                                          const double* __restrict__ B, const double* __restrict__ LR,
                                          const double* __restrict__ LfL, const double* __restrict__ LfR,
                                          double acc[3], double* __restrict__ LBf) {
    const int s = d == 0 ? i : j, o = d == 0 ? j : i;
    const int nT = d == 0 ? P.nxT : P.nyT, oT = d == 0 ? P.nyT : P.nxT;
    if (o < P.gc || o >= oT - P.gc) return false;
    const size_t n = (size_t)P.nxT * P.nyT;
    const double g = P.phys.g;
    const double Dd = d == 0 ? P.dx : P.dy, inv = d == 0 ? P.inv_dx : P.inv_dy;
    const int pow2 = d == 0 ? P.dx_pow2 : P.dy_pow2;
    // total adjoint of cell k's face states: flux part + bottom-source part
    auto face_adj = [&](int kk, int ki, int kj, double aL[3], double aR[3]) {
        const size_t q = cid(P, ki, kj);
        for (int v = 0; v < 3; ++v) { aL[v] = LfL[v * n + q]; aR[v] = LfR[v * n + q]; }
        if (P.bottom && kk >= P.gc && kk < nT - P.gc) {
            const double slope = slope_div(B_face_right(P, B, ki, kj, d) - B_face_left(P, B, ki, kj, d), Dd, inv, pow2);
            const double c = LR[(1 + d) * n + q] * (-g * slope) * 0.5;
            aL[0] += c;
            aR[0] += c;
        }
    };
    double Lc[3], Rc[3];                              // the cell's own face states
    int dry_c = 0;
    // Branch-free over the three stencil offsets, so that every load can be
    // issued up front: an offset whose cell lies outside the range where face
    // states are used is evaluated at a clamped, in-bounds cell (the stencil
    // of a cell in [gc-1, nT-gc] stays inside the array for gc >= 2) and its
    // contribution is dropped.
#pragma unroll
    for (int off = -1; off <= 1; ++off) {
        const int kr = s + off;                       // the reconstructed cell along d
        const bool use = kr >= P.gc - 1 && kr <= nT - P.gc;
        const int kk = kr < P.gc - 1 ? P.gc - 1 : (kr > nT - P.gc ? nT - P.gc : kr);
        const int ki = d == 0 ? kk : i, kj = d == 0 ? j : kk;
        double aL[3], aR[3];
        face_adj(kk, ki, kj, aL, aR);
        double L[3], R[3];
        int pick[3], dry;
        recon_pick(P, U, B, d, ki, kj, L, R, pick, dry);
        const int p = -off;                           // this cell's position in kk's stencil
        const double dc = p == 0 ? 1.0 : 0.0;
        for (int v = 0; v < 3; ++v) {
            const double cs = slope_coef(pick[v], v == 0 ? dry : 0, p, P.theta);
            const double t = aL[v] * (dc - 0.5 * cs) + aR[v] * (dc + 0.5 * cs);
            if (use) acc[v] += t;
        }
        if (off == 0) {
            for (int v = 0; v < 3; ++v) { Lc[v] = L[v]; Rc[v] = R[v]; }
            dry_c = dry;
        }
    }
    const size_t k = cid(P, i, j);
    // the cell's own face bottoms: Bl and Br enter w's face states directly
    // (L_w -= Bl, R_w -= Br) and through the dry-fix line
    if (s >= P.gc - 1 && s <= nT - P.gc) {
        double aL[3], aR[3];
        face_adj(s, i, j, aL, aR);
        const double sbl = dry_c == 1 ? -2.0 : 0.0, sbr = dry_c == 2 ? 2.0 : 0.0;   // d slope_w / d Bl, d Br
        double lBl = aL[0] * ((0.0 - 0.5 * sbl) - 1.0) + aR[0] * (0.0 + 0.5 * sbl);
        double lBr = aL[0] * (0.0 - 0.5 * sbr) + aR[0] * ((0.0 + 0.5 * sbr) - 1.0);
        if (P.bottom && s >= P.gc && s < nT - P.gc) {   // bottom source: -g ((Br - Bl)/Δ) ((hR + hL)/2)
            const double c = LR[(1 + d) * n + k] * (-g) * inv * ((Rc[0] + Lc[0]) / 2.0);
            lBr += c;
            lBl -= c;
        }
        LBf[(2 * d) * n + k] += lBl;
        LBf[(2 * d + 1) * n + k] += lBr;
    }
    return true;
}

// Both directions of the reconstruction adjoint for one cell, one
// read-modify-write of LU: LU += acc_x, then += acc_y (the x and y face-state
// adjoints come in their own buffers).
__global__ void __launch_bounds__(256) k_adj_recon(KParams P, const double* __restrict__ U, const double* __restrict__ B,
                                                   const double* __restrict__ LR, const double* __restrict__ LfL0,
                                                   const double* __restrict__ LfR0, const double* __restrict__ LfL1,
                                                   const double* __restrict__ LfR1, double* __restrict__ LU,
                                                   double* __restrict__ LBf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // the launchers' blocks are 64 wide, so a wave shares threadIdx.y
    const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * blockDim.y + threadIdx.y);
    if (i >= P.nxT || j >= P.nyT) return;
    double ax[3] = {0.0, 0.0, 0.0}, ay[3] = {0.0, 0.0, 0.0};
    const bool hx = recon_dir(P, 0, i, j, U, B, LR, LfL0, LfR0, ax, LBf);
    const bool hy = recon_dir(P, 1, i, j, U, B, LR, LfL1, LfR1, ay, LBf);
    if (!hx && !hy) return;
    const size_t n = (size_t)P.nxT * P.nyT, k = cid(P, i, j);
    for (int v = 0; v < 3; ++v) {
        double lu = LU[v * n + k];
        if (hx) lu += ax[v];
        if (hy) lu += ay[v];
        LU[v * n + k] = lu;
    }
}

// The CFL dt differentiated through its maximum wave speed (simulator.jl:147-150).
// The maximum is maximum(wavespeeds) over cells in memory order, each cell's
// wavespeed being max(right face speed, left face speed) (numericalflux.jl:29-32);
// with equal speeds on several faces the derivative depends on which one the
// maximum returns, so the same one is taken: the first cell in memory order
// whose wavespeed is the maximum (found by the stage-1 face kernel: best =
// its memory index), and its right face unless the left one is strictly
// faster.
__device__ inline void face_speeds(const KParams& P, const double* U, const double* B, int d, int i, int j, double& sr,
                                   double& sl) {
    const int di = d == 0 ? 1 : 0, dj = d == 0 ? 0 : 1;
    double Lm[3], Rm[3], Ll[3], Rl[3], Lr[3], Rr[3], F[3];
    recon_at<double>(P, U, B, d, i, j, Lm, Rm);
    recon_at<double>(P, U, B, d, i - di, j - dj, Ll, Rl);
    recon_at<double>(P, U, B, d, i + di, j + dj, Lr, Rr);
    sr = ad::central_upwind<double>(d, Rm, Lr, P.phys, F);
    sl = ad::central_upwind<double>(d, Rl, Lm, P.phys, F);
}



// Once per reverse step, after the stage tails and both
// directions' face adjoints of stage 1: (1) sum the two stages' block
// partials of the dt and Cz adjoints in a fixed order (sums[0] = the step's
// dt adjoint, cz_total += its Cz adjoint); (2) a step whose dt came from a
// finite positive wave speed must have a face reproducing that speed, else
// the sticky flag is raised; (3) the dt adjoint goes to the face that set dt
// (in the face buffers of its direction, d = 0: LfL0/LfR0, d = 1: LfL1/LfR1).
__global__ void __launch_bounds__(256) k_adj_dtstep(KParams P, const StepCtl* ctl, const double* __restrict__ part_dt,
                                                    const double* __restrict__ part_cz, int count,
                                                    double* __restrict__ sums, double* __restrict__ cz_total,
                                                    const unsigned long long* __restrict__ best, int* __restrict__ flag,
                                                    const double* __restrict__ U, const double* __restrict__ B,
                                                    double* __restrict__ LfL0, double* __restrict__ LfR0,
                                                    double* __restrict__ LfL1, double* __restrict__ LfR1) {
    __shared__ double red[2][256];
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < count; k += blockDim.x) { a += part_dt[k]; b += part_cz[k]; }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { red[0][threadIdx.x] += red[0][threadIdx.x + s]; red[1][threadIdx.x] += red[1][threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double ldt = red[0][0];
    sums[0] = ldt;
    *cz_total += red[1][0];
    const unsigned long long kbest = *best;
    {
        const double ax = __longlong_as_double((long long)ctl->slot[0].wsmax[0]);
        const double ay = __longlong_as_double((long long)ctl->slot[0].wsmax[1]);
        const bool armed = __builtin_isfinite(ax) && __builtin_isfinite(ay) && (ax > 0.0 || ay > 0.0);
        if (armed && kbest == ~0ull) *flag = 1;
    }
    const int d = dt_direction(P, ctl);
    if (kbest >= (unsigned long long)P.nxT * P.nyT) return;
    const int i = (int)(kbest % P.nxT), j = (int)(kbest / P.nxT);
    const double ax = __longlong_as_double((long long)ctl->slot[0].wsmax[d]);
    if (!(ax > 0.0) || !__builtin_isfinite(ax)) return;
    double sr, sl;
    face_speeds(P, U, B, d, i, j, sr, sl);
    const bool right = !(sl > sr);                 // jl_max(sr, sl) returns sr unless sl is larger
    const int di = d == 0 ? 1 : 0, dj = d == 0 ? 0 : 1;
    const int ia = right ? i : i - di, ja = right ? j : j - dj;   // the face lies between (ia,ja) and the next cell
    const int ib = ia + di, jb = ja + dj;
    const double Dd = d == 0 ? P.dx : P.dy;
    // dt = cfl * Δ_d / a_d on the selected direction: d dt / d a = -cfl Δ / a^2
    const double lspeed = ldt * P.cfl * (-Dd / (ax * ax));
    const size_t n = (size_t)P.nxT * P.nyT;
    double La[3], Ra[3], Lb[3], Rb[3];
    recon_at<double>(P, U, B, d, ia, ja, La, Ra);
    recon_at<double>(P, U, B, d, ib, jb, Lb, Rb);
    const double lF0[3] = {0.0, 0.0, 0.0};
    double gm[3], gq[3], sp;
    ad::face_vjp(d, Ra, Lb, P.phys, lF0, gm, gq, sp, lspeed);
    double* LfL = d == 0 ? LfL0 : LfL1;
    double* LfR = d == 0 ? LfR0 : LfR1;
    for (int c = 0; c < 3; ++c) {
        LfR[c * n + cid(P, ia, ja)] += gm[c];
        LfL[c * n + cid(P, ib, jb)] += gq[c];
    }
}

// J = sum (w - B_cell - h_target)^2 over interior cells: block partials, the
// final-state adjoint (w only) and the B_cell adjoint of J itself.
__global__ void __launch_bounds__(256) k_adj_objective(KParams P, const double* __restrict__ U, const double* __restrict__ B,
                                                       const double* __restrict__ target, double* __restrict__ L,
                                                       double* __restrict__ LBc, double* __restrict__ part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // the launchers' blocks are 64 wide, so a wave shares threadIdx.y
    const int j = __builtin_amdgcn_readfirstlane(blockIdx.y * blockDim.y + threadIdx.y);
    double e2 = 0.0;
    if (i >= P.gc && i < P.nxT - P.gc && j >= P.gc && j < P.nyT - P.gc) {
        const size_t k = cid(P, i, j);
        const double h = U[k] - Bc4(P, B, i, j);
        const double e = h - target[(size_t)(i - P.gc) + (size_t)P.nx * (j - P.gc)];
        e2 = e * e;
        L[k] = 2.0 * e;
        LBc[k] += -2.0 * e;
    }
    block_sum(e2, part, blockIdx.x + gridDim.x * blockIdx.y);
}

// Corner gather: B_cell weights 0.25 on its 4 corners, face bottoms 0.5 on 2.
__global__ void k_adj_corners(KParams P, const double* __restrict__ LBc, const double* __restrict__ LBf,
                              double* __restrict__ LB) {
    const int ci = blockIdx.x * blockDim.x + threadIdx.x;
    const int cj = blockIdx.y * blockDim.y + threadIdx.y;
    if (ci > P.nxT || cj > P.nyT) return;
    const size_t n = (size_t)P.nxT * P.nyT;
    auto in = [&](int i, int j) { return i >= 0 && j >= 0 && i < P.nxT && j < P.nyT; };
    double a = 0.0;
    for (int dj = -1; dj <= 0; ++dj)
        for (int di = -1; di <= 0; ++di)
            if (in(ci + di, cj + dj)) a += 0.25 * LBc[cid(P, ci + di, cj + dj)];
    // x faces: left face of (ci, cj) and (ci, cj-1); right face of (ci-1, cj) and (ci-1, cj-1)
    for (int dj = -1; dj <= 0; ++dj) {
        if (in(ci, cj + dj)) a += 0.5 * LBf[0 * n + cid(P, ci, cj + dj)];
        if (in(ci - 1, cj + dj)) a += 0.5 * LBf[1 * n + cid(P, ci - 1, cj + dj)];
    }
    // y faces: bottom face of (ci, cj) and (ci-1, cj); top face of (ci, cj-1) and (ci-1, cj-1)
    for (int di = -1; di <= 0; ++di) {
        if (in(ci + di, cj)) a += 0.5 * LBf[2 * n + cid(P, ci + di, cj)];
        if (in(ci + di, cj - 1)) a += 0.5 * LBf[3 * n + cid(P, ci + di, cj - 1)];
    }
    LB[(size_t)ci + (size_t)(P.nxT + 1) * cj] = a;
}

}  // namespace ad
}  // namespace sfv

// ------------------------------------------------------------------- host
namespace {

using namespace sfv;

// Every buffer starts at zero: the clear is queued on the handle's stream,
// ahead of every kernel of the call (the side stream's work waits for an
// event recorded on that stream after the forward pass).
struct AdjBuffers {
    hipStream_t stream;
    std::vector<void*> owned;
    explicit AdjBuffers(hipStream_t s) : stream(s) {}
    double* alloc(size_t count, std::string& why, int& rc) {
        void* p = nullptr;
        if (rc) return nullptr;
        if (hipMalloc(&p, count * sizeof(double)) != hipSuccess) {
            rc = SFV_ERR_OUT_OF_MEMORY;
            why = "hipMalloc failed for the adjoint buffers";
            return nullptr;
        }
        owned.push_back(p);
        if (hipMemsetAsync(p, 0, count * sizeof(double), stream) != hipSuccess) {
            rc = SFV_ERR_HIP;
            why = "clearing the adjoint buffers failed";
            return nullptr;
        }
        return (double*)p;
    }
    ~AdjBuffers() {
        for (void* p : owned) (void)hipFree(p);
    }
};

// A second stream for the recompute and its events; drained and released
// on every return path.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fwd = nullptr, ready[2] = {nullptr, nullptr}, done[2] = {nullptr, nullptr};
    hipError_t create() {
        hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&fwd, hipEventDisableTiming);
        for (int k = 0; k < 2 && e == hipSuccess; ++k) {
            e = hipEventCreateWithFlags(&ready[k], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&done[k], hipEventDisableTiming);
        }
        return e;
    }
    ~SideStream() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        for (hipEvent_t e : {fwd, ready[0], ready[1], done[0], done[1]})
            if (e) (void)hipEventDestroy(e);
    }
};

#define ATRY(expr)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) {                                                                      \
            h->err = std::string(#expr) + ": " + hipGetErrorString(e_);                              \
            return SFV_ERR_HIP;                                                                      \
        }                                                                                            \
    } while (0)

// One step U0 -> U2 of the forward pass with the march kernels, in step
// slot `slot` of a loop prepared once (stage 2 advances t and arms the next
// slot; U2's ghost cells are left unwritten and filled where a state leaves
// the march path).  The adjoint accepts only handles the march path serves.
int march_step(sfv_handle* h, int slot, const double* U0, double* U2) {
    const KParams& P = h->P;
    hipStream_t st = h->stream;
    ATRY(launch_march<double>(st, P, h->ctl, slot, 1, U0, nullptr, (double*)h->RHS, h->B, 0, 1));
    ATRY(launch_march<double>(st, P, h->ctl, slot, 2, U0, (const double*)h->RHS, U2, h->B, 1));
    return SFV_OK;
}

// One forward step U0 -> U2 (bitwise the fused path), keeping the stage-1
// state, both right-hand sides at every cell and, in ctl_out, the step's
// control block as stage 1 leaves it (its dt and wave-speed maxima).  The
// right-hand sides come from the march stage-1 kernel, which reads the
// ghost cells the boundary condition filled and writes interior cells; the
// finishing kernel stores the ghost cells' (sources only) beside them.  The
// second march launch raises the slot's maxima again, after the copy.
int forward_step(sfv_handle* h, hipStream_t st, double t, const double* U0, double* U1, double* R1, double* R2,
                 double* U2, StepCtl* ctl_out) {
    const KParams& P = h->P;
    ATRY(launch_prepare(st, P, h->ctl, 0, t, NAN, 0, 0, 0.0, 1, 0, SFV_F64));
    ATRY(launch_march<double>(st, P, h->ctl, 0, 1, U0, nullptr, R1, h->B));
    ATRY(launch_finish<double>(st, P, h->ctl, 0, 1, U0, U0, R1, U1, h->B, 1, 0, R1));
    ATRY(hipMemcpyAsync(ctl_out, h->ctl, sizeof(StepCtl), hipMemcpyDeviceToDevice, st));
    ATRY(launch_bc<double>(st, P, h->ctl, 0, U1, BC_IN_STEP));
    ATRY(launch_march<double>(st, P, h->ctl, 0, 1, U1, nullptr, R2, h->B));
    ATRY(launch_finish<double>(st, P, h->ctl, 0, 2, U0, U1, R2, U2, h->B, 1, 0, R2));
    ATRY(launch_bc<double>(st, P, h->ctl, 0, U2, BC_LAST_STAGE));
    return SFV_OK;
}

}  // namespace

extern "C" int sfv_adjoint(sfv_handle* h, int64_t nsteps, int32_t checkpoint_every, double* t_io,
                           const double* h_target, double* J_out, double* dJ_dB, double* dJ_dCz, double* dJ_dU0) {
    if (!h || !t_io || !h_target) return SFV_ERR_ARGUMENT;
    const sfv_config& c = h->cfg;
    if (c.precision != SFV_F64 || c.stepper != SFV_RUNGE_KUTTA2 || c.reconstruction != SFV_LINEAR_RECONSTRUCTION ||
        c.rain == SFV_RAIN_FIELD || c.embedded_1d || !h->fusable()) {
        h->err = "the adjoint runs Float64 RungeKutta2 with LinearReconstruction, ghost sources inside the grid "
                 "(WallBC or PeriodicBC), and no FunctionalRain";
        return SFV_ERR_ARGUMENT;
    }
    if (nsteps < 1 || checkpoint_every < 1) { h->err = "nsteps and checkpoint_every must be >= 1"; return SFV_ERR_ARGUMENT; }
    if (!h->has_state) { h->err = "set the state before the adjoint"; return SFV_ERR_STATE; }
    ATRY(hipSetDevice(h->device));
    const KParams& P = h->P;
    hipStream_t st = h->stream;
    const size_t n = h->ncell, n3 = 3 * n, nB = (size_t)(P.nxT + 1) * (P.nyT + 1);
    const int64_t ce = checkpoint_every;
    const int64_t nseg = (nsteps + ce - 1) / ce;
    std::string why;
    int rc = SFV_OK;
    AdjBuffers buf(st);
    double* ck = buf.alloc(n3 * (size_t)nseg, why, rc);         // checkpoint states
    double* seg = buf.alloc(n3 * (size_t)(ce + 1), why, rc);    // the states of one segment
    double* U1 = buf.alloc(n3, why, rc);
    double* R1 = buf.alloc(n3, why, rc);
    double* R2 = buf.alloc(n3, why, rc);
    double* Ut = buf.alloc(n3, why, rc);                        // scratch output
    double* L = buf.alloc(n3, why, rc);                         // adjoint of the current step's output
    double* Lnew = buf.alloc(n3, why, rc);                      // adjoint of its input
    double* L1 = buf.alloc(n3, why, rc);                        // adjoint of U1
    double* LR = buf.alloc(n3, why, rc);
    double* LfL = buf.alloc(n3, why, rc);                       // x face-state adjoints
    double* LfR = buf.alloc(n3, why, rc);
    double* LfL1 = buf.alloc(n3, why, rc);                      // y face-state adjoints
    double* LfR1 = buf.alloc(n3, why, rc);
    double* LBc = buf.alloc(n, why, rc);
    double* LBf = buf.alloc(4 * n, why, rc);
    double* LB = buf.alloc(nB, why, rc);
    double* tgt = buf.alloc((size_t)P.nx * P.ny, why, rc);
    const dim3 cb(64, 4), cg((P.nxT + 63) / 64, (P.nyT + 3) / 4);
    const int nblk = (int)(cg.x * cg.y);
    double* part_dt = buf.alloc(2 * (size_t)nblk, why, rc);
    double* part_cz = buf.alloc(2 * (size_t)nblk, why, rc);
    double* sums = buf.alloc(4, why, rc);                      // [0] dt adjoint of the step, [1] Cz total, [2] J
    double* part_j = buf.alloc((size_t)nblk, why, rc);
    unsigned long long* best = (unsigned long long*)buf.alloc(1, why, rc);
    double* d_times = buf.alloc((size_t)nsteps + 1, why, rc);
    int* flag = (int*)buf.alloc(1, why, rc);
    // the control block of the step being recomputed, without kept blocks
    StepCtl* step_ctl = (StepCtl*)buf.alloc(sizeof(StepCtl) / sizeof(double) + 1, why, rc);
    // The stage-1 states, right-hand sides and step controls of a segment are
    // kept from its recompute when they fit in half the free device memory;
    // otherwise each reverse step recomputes its own.
    size_t free_b = 0, total_b = 0;
    const size_t keep_bytes = (size_t)ce * (3 * n3 * sizeof(double) + sizeof(StepCtl));
    const char* keep_env = std::getenv("SFV_ADJOINT_KEEP");      // "0": always recompute
    const bool keep = !rc && !(keep_env && keep_env[0] == '0') && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
                      keep_bytes < free_b / 2;
    // When two sets of them (and a second segment of states) fit as well,
    // segment sg-1 is recomputed on a second stream while the reverse steps
    // of segment sg run.
    const char* ovl_env = std::getenv("SFV_ADJOINT_OVERLAP");   // "0": one stream
    const size_t seg_bytes = n3 * (size_t)(ce + 1) * sizeof(double);
    const bool overlap = keep && nseg > 1 && !(ovl_env && ovl_env[0] == '0') &&
                         2 * keep_bytes + seg_bytes < free_b / 2;
    const size_t ctl_words = (size_t)ce * (sizeof(StepCtl) / sizeof(double) + 1);
    double* kU1s[2] = {nullptr, nullptr};
    double* kR1s[2] = {nullptr, nullptr};
    double* kR2s[2] = {nullptr, nullptr};
    StepCtl* kctls[2] = {nullptr, nullptr};
    double* segs[2] = {seg, nullptr};
    for (int k = 0; k < (overlap ? 2 : keep ? 1 : 0); ++k) {
        kU1s[k] = buf.alloc(n3 * (size_t)ce, why, rc);
        kR1s[k] = buf.alloc(n3 * (size_t)ce, why, rc);
        kR2s[k] = buf.alloc(n3 * (size_t)ce, why, rc);
        kctls[k] = (StepCtl*)buf.alloc(ctl_words, why, rc);
    }
    if (overlap) segs[1] = buf.alloc(n3 * (size_t)(ce + 1), why, rc);
    if (rc) { h->err = why; return rc; }
    ATRY(hipMemcpyAsync(tgt, h_target, sizeof(double) * P.nx * P.ny, hipMemcpyHostToDevice, st));

    // forward pass with checkpoints, device-resident: one prepared loop of
    // march steps; each step's end time is copied on the device and the
    // times come back once
    std::vector<double> times((size_t)nsteps + 1);
    double* cur = (double*)h->U[h->cur & 1];
    ATRY(hipMemcpyAsync(seg, cur, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    ATRY(launch_prepare(st, P, h->ctl, 0, *t_io, NAN, 0, 0, 0.0, -1, 0, SFV_F64));
    for (int64_t s = 0; s < nsteps; ++s) {
        double* a = seg + (size_t)(s & 1) * n3;
        double* b = seg + (size_t)((s + 1) & 1) * n3;
        if (s % ce == 0) {
            // a checkpoint restarts the per-pass recompute, which reads ghosts
            if (s > 0) ATRY(launch_bc<double>(st, P, h->ctl, 0, a, BC_STANDALONE));
            ATRY(hipMemcpyAsync(ck + (size_t)(s / ce) * n3, a, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        if ((rc = march_step(h, (int)(s & 1), a, b))) return rc;
        ATRY(hipMemcpyAsync(d_times + s + 1, &h->ctl->t, sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    ATRY(hipMemcpyAsync(times.data() + 1, d_times + 1, (size_t)nsteps * sizeof(double), hipMemcpyDeviceToHost, st));
    ATRY(hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, st));
    ATRY(hipStreamSynchronize(st));
    times[0] = *t_io;
    const double t = times[(size_t)nsteps];
    const double last_dt = h->ctl_host->dt;
    double* UT = seg + (size_t)(nsteps & 1) * n3;
    // the final state, ghosts filled, becomes the handle's state
    if (nsteps > 0) ATRY(launch_bc<double>(st, P, h->ctl, 0, UT, BC_STANDALONE));
    ATRY(hipMemcpyAsync(cur, UT, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));

    // objective and its adjoint
    hipLaunchKernelGGL(ad::k_adj_objective, cg, cb, 0, st, P, UT, h->B, tgt, L, LBc, part_j);
    ATRY(hipPeekAtLastError());
    std::vector<double> pj((size_t)nblk);
    ATRY(hipMemcpyAsync(pj.data(), part_j, sizeof(double) * nblk, hipMemcpyDeviceToHost, st));
    ATRY(hipStreamSynchronize(st));
    double J = 0.0;
    for (double v : pj) J += v;

    const int bcx = (P.ny + 255) / 256, bcy = (P.nx + 4 * P.gc * P.gc + 255) / 256;
    auto faces_grid = [&](int d) {
        const int nT = d == 0 ? P.nxT : P.nyT, nO = d == 0 ? P.ny : P.nx;
        return dim3((unsigned)(((nT - 2 * P.gc + 1) * nO + 255) / 256));
    };
    // Recompute of segment sg from its checkpoint into buffer set `set`, on
    // stream rs: with keep, every step's stage-1 state, right-hand sides and
    // control block are kept; without, only the states.
    auto recompute = [&](int64_t sg, int set, hipStream_t rs) -> int {
        const int64_t s0 = sg * ce, s1 = std::min<int64_t>(nsteps, s0 + ce);
        double* sb = segs[set];
        ATRY(hipMemcpyAsync(sb, ck + (size_t)sg * n3, n3 * sizeof(double), hipMemcpyDeviceToDevice, rs));
        for (int64_t s = s0; s < (keep ? s1 : s1 - 1); ++s) {
            double* a = sb + (size_t)(s - s0) * n3;
            const size_t o = (size_t)(s - s0) * n3;
            int r;
            if (keep) {
                if ((r = forward_step(h, rs, times[(size_t)s], a, kU1s[set] + o, kR1s[set] + o, kR2s[set] + o,
                                      a + n3, kctls[set] + (s - s0))))
                    return r;
            } else if ((r = forward_step(h, rs, times[(size_t)s], a, U1, R1, R2, a + n3, step_ctl))) {
                return r;
            }
        }
        return SFV_OK;
    };
    // The reverse steps of segment sg (buffer set `set`), on the main stream.
    auto reverse = [&](int64_t sg, int set) -> int {
        const int64_t s0 = sg * ce, s1 = std::min<int64_t>(nsteps, s0 + ce);
        for (int64_t s = s1 - 1; s >= s0; --s) {
            const size_t o = (size_t)(s - s0) * n3;
            const double* U0 = segs[set] + o;
            const double *u1 = U1, *r1 = R1, *r2 = R2;
            const StepCtl* cs = step_ctl;               // the step's control block (its dt and maxima)
            if (keep) {
                u1 = kU1s[set] + o;
                r1 = kR1s[set] + o;
                r2 = kR2s[set] + o;
                cs = kctls[set] + (s - s0);
            } else {
                const int r = forward_step(h, st, times[(size_t)s], U0, U1, R1, R2, Ut, step_ctl);
                if (r) return r;
            }
            // stage 2: U2 = clamp(friction(0.5 (U0 + U1 + dt R2))); the tail
            // reads L through the boundary condition's fold and writes L1 and
            // Lnew afresh
            hipLaunchKernelGGL(ad::k_adj_tail, cg, cb, 0, st, P, cs, 2, U0, u1, r2, h->B, L, L1, Lnew, LR, LBc,
                               part_dt, part_cz, best);
            hipLaunchKernelGGL(ad::k_adj_faces, faces_grid(0), dim3(256), 0, st, P, 0, u1, h->B, LR, LfL, LfR, cs,
                               nullptr);
            hipLaunchKernelGGL(ad::k_adj_faces, faces_grid(1), dim3(256), 0, st, P, 1, u1, h->B, LR, LfL1, LfR1, cs,
                               nullptr);
            hipLaunchKernelGGL(ad::k_adj_recon, cg, cb, 0, st, P, u1, h->B, LR, LfL, LfR, LfL1, LfR1, L1, LBf);
            // stage 1: U1 = clamp(friction(U0 + dt R1))
            hipLaunchKernelGGL(ad::k_adj_tail, cg, cb, 0, st, P, cs, 1, U0, U0, r1, h->B, L1, Lnew, nullptr, LR,
                               LBc, part_dt, part_cz, best);
            // stage 1's faces also find the face that set dt, as the
            // reference's maximum(ws) would
            hipLaunchKernelGGL(ad::k_adj_faces, faces_grid(0), dim3(256), 0, st, P, 0, U0, h->B, LR, LfL, LfR, cs,
                               best);
            hipLaunchKernelGGL(ad::k_adj_faces, faces_grid(1), dim3(256), 0, st, P, 1, U0, h->B, LR, LfL1, LfR1, cs,
                               best);
            hipLaunchKernelGGL(ad::k_adj_dtstep, dim3(1), dim3(256), 0, st, P, cs, part_dt, part_cz, 2 * nblk,
                               sums, sums + 1, best, flag, U0, h->B, LfL, LfR, LfL1, LfR1);
            hipLaunchKernelGGL(ad::k_adj_recon, cg, cb, 0, st, P, U0, h->B, LR, LfL, LfR, LfL1, LfR1, Lnew, LBf);
            ATRY(hipPeekAtLastError());
            std::swap(L, Lnew);
        }
        return SFV_OK;
    };
    SideStream side;
    if (overlap) {
        // the side stream starts after the forward pass; segment sg's reverse
        // waits for its recompute, and a recompute into a buffer set waits
        // for the reverse that last read it
        ATRY(side.create());
        ATRY(hipEventRecord(side.fwd, st));
        ATRY(hipStreamWaitEvent(side.s, side.fwd, 0));
        bool used[2] = {false, false};
        const int last = (int)((nseg - 1) & 1);
        if ((rc = recompute(nseg - 1, last, side.s))) return rc;
        ATRY(hipEventRecord(side.ready[last], side.s));
        for (int64_t sg = nseg - 1; sg >= 0; --sg) {
            const int set = (int)(sg & 1), other = 1 - set;
            ATRY(hipStreamWaitEvent(st, side.ready[set], 0));
            if (sg > 0) {
                if (used[other]) ATRY(hipStreamWaitEvent(side.s, side.done[other], 0));
                if ((rc = recompute(sg - 1, other, side.s))) return rc;
                ATRY(hipEventRecord(side.ready[other], side.s));
            }
            if ((rc = reverse(sg, set))) return rc;
            ATRY(hipEventRecord(side.done[set], st));
            used[set] = true;
        }
    } else {
        for (int64_t sg = nseg - 1; sg >= 0; --sg) {
            if ((rc = recompute(sg, 0, st))) return rc;
            if ((rc = reverse(sg, 0))) return rc;
        }
    }
    // the initial state's own boundary condition
    hipLaunchKernelGGL(ad::k_adj_bc, dim3(bcx), dim3(256), 0, st, P, L, 0);
    hipLaunchKernelGGL(ad::k_adj_bc, dim3(bcy), dim3(256), 0, st, P, L, 1);
    hipLaunchKernelGGL(ad::k_adj_corners, dim3((P.nxT + 1 + 63) / 64, (P.nyT + 1 + 3) / 4), cb, 0, st, P, LBc, LBf, LB);
    ATRY(hipPeekAtLastError());
    double czsum = 0.0;
    int missing = 0;
    ATRY(hipMemcpyAsync(&missing, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    ATRY(hipMemcpyAsync(&czsum, sums + 1, sizeof(double), hipMemcpyDeviceToHost, st));
    if (dJ_dB) ATRY(hipMemcpyAsync(dJ_dB, LB, nB * sizeof(double), hipMemcpyDeviceToHost, st));
    std::vector<double> Lh;
    if (dJ_dU0) {
        Lh.resize(n3);
        ATRY(hipMemcpyAsync(Lh.data(), L, n3 * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    ATRY(hipStreamSynchronize(st));
    if (missing) {
        h->err = "adjoint: no face reproduces a step's maximum wave speed";
        return SFV_ERR_STATE;
    }
    if (dJ_dU0)
        for (int v = 0; v < 3; ++v)
            for (int j = 0; j < P.ny; ++j)
                for (int i = 0; i < P.nx; ++i)
                    dJ_dU0[(size_t)i + (size_t)P.nx * (j + (size_t)P.ny * v)] =
                        Lh[v * n + (size_t)(i + P.gc) + (size_t)P.nxT * (j + P.gc)];
    if (dJ_dCz) *dJ_dCz = czsum;
    if (J_out) *J_out = J;
    *t_io = t;
    h->last_dt = last_dt;
    h->have_dt = true;
    return SFV_OK;
}
