// swe_physics.hpp — per-cell and per-face arithmetic of the practical 2D
// shallow-water scheme, as device functions for gfx950.
//
// Every expression keeps the reference's operation order and rounding points
// so that FP64 results are bitwise equal to the reference's CPU path; the
// library is built with -ffp-contract=off (no a*b+c fusion) and without
// fast-math (IEEE NaN / signed-zero behaviour).  Storage type S is double
// (Float64 backend) or float (Float32 storage); values read from S buffers
// stay S and parameters stay double, so C++ promotion reproduces Julia's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfv {

// Julia max/min on IEEE floats (base/math.jl): NaN-propagating, -0.0 < +0.0.
template <class T>
__device__ __forceinline__ T jl_max(T x, T y) {
    if (x != x) return x;
    if (y != y) return y;
    if (x > y) return x;
    if (y > x) return y;
    return __builtin_signbit(x) ? y : x;
}
template <class T>
__device__ __forceinline__ T jl_min(T x, T y) {
    if (x != x) return x;
    if (y != y) return y;
    if (x < y) return x;
    if (y < x) return y;
    return __builtin_signbit(x) ? x : y;
}

// linearreconstruction.jl:28-35
__device__ __forceinline__ double minmod(double a, double b, double c) {
    if ((a > 0) && (b > 0) && (c > 0)) return jl_min(jl_min(a, b), c);
    if (a < 0 && b < 0 && c < 0) return jl_max(jl_max(a, b), c);
    return 0.0;
}

// equation.jl:56 (Brodtkorb & Holm 2021, eq. 26).  When |h| >= κ the inner
// min(...) is <= κ <= |h|, so max(|h|, ...) = |h| and h* = copysign(1,h)*|h|
// = h exactly; the divide is only evaluated where it can matter.
template <class H>
__device__ __forceinline__ double desingularize(H h, double kappa) {
    const double ah = (double)__builtin_fabs(h);
    if (ah >= kappa) return (double)h;
    const H h2 = h * h;
    const double sgn = __builtin_signbit(h) ? -1.0 : 1.0;
    return sgn * jl_max(ah, jl_min(h2 / (2 * kappa) + kappa / 2.0, kappa));
}

struct PhysParams {
    double g, rho, cutoff, kappa;
};

// swe_2D.jl:43-81: flux and eigenvalues of one wet face state in direction d
// (0 = x, 1 = y).
template <class S>
__device__ __forceinline__ void phys_flux(int d, S h, S hu, S hv, const PhysParams& p,
                                          double f[3], double e[2]) {
    const double hs = desingularize(h, p.kappa);
    const double u = hu / hs;
    const double v = hv / hs;
    const S h2 = h * h;
    const double c = sqrt(p.g * h);
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

// centralupwind.jl:48-77 (practical SWE variant).  m = faceminus (right face
// of the left cell), q = faceplus (left face of the right cell).  Returns the
// face wave speed; F receives the numerical flux.
template <class S>
__device__ __forceinline__ double central_upwind(int d, const S m[3], const S q[3], const PhysParams& p,
                                                 double F[3]) {
    double fm[3] = {0.0, 0.0, 0.0}, em[2] = {0.0, 0.0};
    double fp[3] = {0.0, 0.0, 0.0}, ep[2] = {0.0, 0.0};
    if (m[0] > p.cutoff) phys_flux(d, m[0], m[1], m[2], p, fm, em);
    if (q[0] > p.cutoff) phys_flux(d, q[0], q[1], q[2], p, fp, ep);
    const double aplus = jl_max(jl_max(ep[0], em[0]), 0.0);
    const double aminus = jl_min(jl_min(ep[1], em[1]), 0.0);
    const double da = aplus - aminus;
    if (__builtin_fabs(da) < p.kappa) {
        F[0] = F[1] = F[2] = 0.0;
        return 0.0;
    }
    const double pm = (aplus * aminus) / da;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const S diff = q[v] - m[v];
        F[v] = (aplus * fm[v] - aminus * fp[v]) / da + pm * diff;
    }
    if (m[0] < p.cutoff && q[0] < p.cutoff) return 0.0;
    return jl_max(__builtin_fabs(aplus), __builtin_fabs(aminus));
}

// minmod_slope + the SWE dry-face slope fix + face values
// (linearreconstruction.jl:37-43, 73-100).  Given the cell and its two
// neighbours along the sweep, and the cell's two face bottoms, writes the
// left and right face states (h, hu, hv) — variable 1 becomes a depth.
template <class S>
__device__ __forceinline__ void reconstruct_cell(const S l[3], const S c[3], const S r[3], double theta,
                                                 double Bl, double Br, S L[3], S R[3]) {
    double slope[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const S fwd = r[v] - c[v];
        const S bwd = c[v] - l[v];
        const double central = (fwd + bwd) / 2.0;
        slope[v] = minmod(theta * fwd, central, theta * bwd);
    }
    const S w = c[0];
    if (w - 0.5 * slope[0] < Bl) slope[0] = 2.0 * (w - Bl);
    else if (w + 0.5 * slope[0] < Br) slope[0] = 2.0 * (Br - w);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        L[v] = (S)(c[v] - 0.5 * slope[v]);
        R[v] = (S)(c[v] + 0.5 * slope[v]);
    }
    L[0] = (S)(L[0] - Bl);
    R[0] = (S)(R[0] - Br);
}

// friction.jl:30-43
__device__ __forceinline__ double friction_law(int law, double c, double h, double speed) {
    if (law == 2) { const double denom = cbrt(h) * h; return -c * speed / denom; }
    if (law == 3) { const double denom = cbrt(h) * h * h; return -c * speed / denom; }
    const double denom = h * h;
    return -c * speed / denom;
}

// friction.jl:62-78: the implicit friction scalar from the substep input
// state (w, hu, hv) and B_cell.
template <class S>
__device__ __forceinline__ double friction_scalar(int law, double Cz, double kappa, S w, S hu, S hv, double Bm) {
    const double hs = desingularize(w - Bm, kappa);
    const double u = hu / hs, v = hv / hs;
    double speed = sqrt(u * u + v * v);
    if (speed == 0.0) speed = 0.0;
    return friction_law(law, Cz, hs, speed);
}

}  // namespace sfv
