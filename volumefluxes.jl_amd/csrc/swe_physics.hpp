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

// max / min of two values neither of which is -0.0 (wave speeds and
// eigenvalues: u +- c with c > 0 rounds to +0.0 at worst, dry sides give
// +0.0): one hardware max/min, and Julia's NaN propagation kept by an
// unordered check.  For such inputs this equals jl_max / jl_min.
__device__ __forceinline__ double nz_max(double a, double b) { return (a > b || a != a) ? a : b; }
__device__ __forceinline__ double nz_min(double a, double b) { return (a < b || a != a) ? a : b; }

// linearreconstruction.jl:28-35: min(a,b,c) if all > 0, max(a,b,c) if all
// < 0, else +0.0 (also when any is NaN, as every comparison is then false).
// Evaluated as max(lo, 0) + min(hi, 0) (+0.0) with lo/hi the hardware
// min/max: all > 0 gives lo, all < 0 gives hi, mixed signs or zeros give a
// sum of two zeros, and the trailing +0.0 turns -0.0 + -0.0 into +0.0.
// NaN: with `regular` (θ finite and nonzero) a NaN in a or c implies one in b
// (a = θ·fwd, c = θ·bwd, b = (fwd + bwd)/2), so checking b suffices.
__device__ __forceinline__ double minmod(double a, double b, double c, bool regular = false) {
    const double lo = __builtin_fmin(__builtin_fmin(a, b), c);
    const double hi = __builtin_fmax(__builtin_fmax(a, b), c);
    const double r = (__builtin_fmax(lo, 0.0) + __builtin_fmin(hi, 0.0)) + 0.0;
    const bool nan = regular ? (b != b) : (__builtin_isunordered(a, b) || c != c);
    return nan ? 0.0 : r;
}

// equation.jl:56 (Brodtkorb & Holm 2021, eq. 26).  When |h| >= κ the inner
// min(...) is <= κ <= |h|, so max(|h|, ...) = |h| and h* = copysign(1,h)*|h|
// = h exactly; the divide is only evaluated where it can matter.
template <class H>
__device__ __forceinline__ double desingularize(H h, double kappa) {
    const double ah = (double)__builtin_fabs(h);
    if (!(ah < kappa)) return (double)h;          // |h| >= κ, or NaN (which propagates)
    const H h2 = h * h;
    const double sgn = __builtin_signbit(h) ? -1.0 : 1.0;
    // |h| < κ here, and h^2/(2κ) + κ/2 >= κ/2 > 0: no NaN or -0.0 reaches max/min
    return sgn * __builtin_fmax(ah, __builtin_fmin(h2 / (2 * kappa) + kappa / 2.0, kappa));
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
    const double aplus = nz_max(nz_max(ep[0], em[0]), 0.0);
    const double aminus = nz_min(nz_min(ep[1], em[1]), 0.0);
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
    return nz_max(__builtin_fabs(aplus), __builtin_fabs(aminus));
}

// minmod_slope + the SWE dry-face slope fix + face values
// (linearreconstruction.jl:37-43, 73-100).  Given the cell and its two
// neighbours along the sweep, and the cell's two face bottoms, writes the
// left and right face states (h, hu, hv) — variable 1 becomes a depth.
template <class S>
__device__ __forceinline__ void reconstruct_cell(const S l[3], const S c[3], const S r[3], double theta,
                                                 double Bl, double Br, S L[3], S R[3], bool known_regular = false) {
    const bool regular = known_regular || (__builtin_isfinite(theta) && theta != 0.0);
    double slope[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const S fwd = r[v] - c[v];
        const S bwd = c[v] - l[v];
        const double central = (fwd + bwd) / 2.0;
        slope[v] = minmod(theta * fwd, central, theta * bwd, regular);
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

// implicit_friction's `out ./ (1 .- dt .* [0, f, f])` (friction.jl:77).  For
// the w component the divisor 1 - dt*0.0 is exactly 1.0 for finite dt and NaN
// otherwise, so multiplying by it is the same as dividing.
template <class S>
__device__ __forceinline__ void apply_friction(S o[3], S dt_s, double fs) {
    o[0] = (S)(o[0] * (1 - dt_s * 0.0));
    o[1] = (S)(o[1] / (1 - dt_s * fs));
    o[2] = (S)(o[2] / (1 - dt_s * fs));
}

// (B_R - B_L)/Δ of the bottom source (source_term_bottom.jl:38).  When Δ is a
// power of two, x/Δ and x*(1/Δ) are the same correctly rounded value.
__device__ __forceinline__ double slope_div(double x, double d, double inv_d, int pow2) {
    return pow2 ? x * inv_d : x / d;
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


// ---------------------------------------------------------------------------
// Branch-free forms for the fused march kernels.  Each equals the function
// above for every input; the reference's branches become selects so that a
// wave's face work is straight-line code the scheduler can interleave.

// minmod slopes + SWE slope fix + face values, with the fix as a select.
template <class S>
__device__ __forceinline__ void reconstruct_cell_nb(const S l[3], const S c[3], const S r[3], double theta,
                                                    double Bl, double Br, S L[3], S R[3]) {
    const bool regular = __builtin_isfinite(theta) && theta != 0.0;
    double slope[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const S fwd = r[v] - c[v];
        const S bwd = c[v] - l[v];
        const double central = (fwd + bwd) / 2.0;
        slope[v] = minmod(theta * fwd, central, theta * bwd, regular);
    }
    const S w = c[0];
    const bool fix_l = w - 0.5 * slope[0] < Bl;
    const bool fix_r = w + 0.5 * slope[0] < Br;
    const double sl = 2.0 * (w - Bl), sr = 2.0 * (Br - w);
    slope[0] = fix_l ? sl : (fix_r ? sr : slope[0]);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        L[v] = (S)(c[v] - 0.5 * slope[v]);
        R[v] = (S)(c[v] + 0.5 * slope[v]);
    }
    L[0] = (S)(L[0] - Bl);
    R[0] = (S)(R[0] - Br);
}

// Desingularised depths of N face sides: the |h| >= κ case (h* = h) inline,
// the rare |h| < κ case for all of them in one branch.
template <class S, int N>
__device__ __forceinline__ void desingularize_n(const S* const h[N], double kappa, double hs[N]) {
    bool slow = false;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        hs[k] = (double)*h[k];
        slow |= (double)__builtin_fabs(*h[k]) < kappa;
    }
    if (slow) {
#pragma unroll
        for (int k = 0; k < N; ++k)
            if ((double)__builtin_fabs(*h[k]) < kappa) hs[k] = desingularize(*h[k], kappa);
    }
}

// centralupwind.jl:48-77 for direction D with the desingularised depths of
// both sides given: both sides are evaluated and dry sides (h <= cutoff)
// select zero flux and eigenvalues, exactly as the reference's zero(...)
// initialisers; the degenerate |a+ - a-| < κ and both-dry cases select the
// reference's early returns.
template <int D, class S>
__device__ __forceinline__ double central_upwind_nb(const S m[3], const S q[3], double hsm, double hsq,
                                                    const PhysParams& p, double F[3]) {
    const bool wm = m[0] > p.cutoff, wq = q[0] > p.cutoff;
    double fm[3], fp[3], em[2], ep[2];
    {
        const double u = m[1] / hsm, v = m[2] / hsm;
        const S h2 = m[0] * m[0];
        const double c = sqrt(p.g * m[0]);
        const double a = D == 0 ? u : v;
        const double f0 = p.rho * m[0] * a;
        const double f1 = D == 0 ? p.rho * m[0] * u * u + 0.5 * p.rho * p.g * h2 : p.rho * m[0] * u * v;
        const double f2 = D == 0 ? p.rho * m[0] * u * v : p.rho * m[0] * v * v + 0.5 * p.rho * p.g * h2;
        fm[0] = wm ? f0 : 0.0;
        fm[1] = wm ? f1 : 0.0;
        fm[2] = wm ? f2 : 0.0;
        em[0] = wm ? a + c : 0.0;
        em[1] = wm ? a - c : 0.0;
    }
    {
        const double u = q[1] / hsq, v = q[2] / hsq;
        const S h2 = q[0] * q[0];
        const double c = sqrt(p.g * q[0]);
        const double a = D == 0 ? u : v;
        const double f0 = p.rho * q[0] * a;
        const double f1 = D == 0 ? p.rho * q[0] * u * u + 0.5 * p.rho * p.g * h2 : p.rho * q[0] * u * v;
        const double f2 = D == 0 ? p.rho * q[0] * u * v : p.rho * q[0] * v * v + 0.5 * p.rho * p.g * h2;
        fp[0] = wq ? f0 : 0.0;
        fp[1] = wq ? f1 : 0.0;
        fp[2] = wq ? f2 : 0.0;
        ep[0] = wq ? a + c : 0.0;
        ep[1] = wq ? a - c : 0.0;
    }
    const double aplus = nz_max(nz_max(ep[0], em[0]), 0.0);
    const double aminus = nz_min(nz_min(ep[1], em[1]), 0.0);
    const double da = aplus - aminus;
    const bool degenerate = __builtin_fabs(da) < p.kappa;
    const double pm = (aplus * aminus) / da;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const S diff = q[v] - m[v];
        const double f = (aplus * fm[v] - aminus * fp[v]) / da + pm * diff;
        F[v] = degenerate ? 0.0 : f;
    }
    const bool both_dry = m[0] < p.cutoff && q[0] < p.cutoff;
    const double sp = nz_max(__builtin_fabs(aplus), __builtin_fabs(aminus));
    return (degenerate || both_dry) ? 0.0 : sp;
}

// centralupwind.jl:48-77 for direction D when both sides are wet and
// |h| >= κ on both (so h* = h, equation.jl:56): the reference's branches are
// all decided, and what is left is straight-line code — the two sides and the
// four quotients by a+ - a- are independent chains.  Only the degenerate
// |a+ - a-| < κ return remains, as a select.
template <int D, class S>
__device__ __forceinline__ double central_upwind_wet(const S m[3], const S q[3], const PhysParams& p, double F[3]) {
    const double um = m[1] / (double)m[0], vm = m[2] / (double)m[0];
    const double uq = q[1] / (double)q[0], vq = q[2] / (double)q[0];
    const S hm2 = m[0] * m[0], hq2 = q[0] * q[0];
    const double cm = sqrt(p.g * m[0]), cq = sqrt(p.g * q[0]);
    const double am = D == 0 ? um : vm, aq = D == 0 ? uq : vq;
    const double rgh = 0.5 * p.rho * p.g;
    double fm[3], fp[3];
    fm[0] = p.rho * m[0] * am;
    fp[0] = p.rho * q[0] * aq;
    if (D == 0) {
        fm[1] = p.rho * m[0] * um * um + rgh * hm2;
        fm[2] = p.rho * m[0] * um * vm;
        fp[1] = p.rho * q[0] * uq * uq + rgh * hq2;
        fp[2] = p.rho * q[0] * uq * vq;
    } else {
        fm[1] = p.rho * m[0] * um * vm;
        fm[2] = p.rho * m[0] * vm * vm + rgh * hm2;
        fp[1] = p.rho * q[0] * uq * vq;
        fp[2] = p.rho * q[0] * vq * vq + rgh * hq2;
    }
    const double aplus = nz_max(nz_max(aq + cq, am + cm), 0.0);
    const double aminus = nz_min(nz_min(aq - cq, am - cm), 0.0);
    const double da = aplus - aminus;
    const bool degenerate = __builtin_fabs(da) < p.kappa;
    const double pm = (aplus * aminus) / da;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const S diff = q[v] - m[v];
        const double f = (aplus * fm[v] - aminus * fp[v]) / da + pm * diff;
        F[v] = degenerate ? 0.0 : f;
    }
    const double sp = nz_max(__builtin_fabs(aplus), __builtin_fabs(aminus));
    return degenerate ? 0.0 : sp;
}

// ---------------------------------------------------------------------------
// Face evaluation on operands in a proven range.
//
// The compiler's correctly rounded FP64 quotient x/d on gfx950 is
//   ds = div_scale(d; d, x), xs = div_scale(x; d, x, flag),
//   r = rcp(ds), then twice r = fma(r, fma(-ds, r, 1), r),
//   q = xs*r, e = fma(-ds, q, xs), q = div_fmas(e, r, q, flag),
//   result = div_fixup(q, d, x).
// div_scale returns its operand unchanged and clears the flag, and div_fmas
// is then a plain fma, unless an operand or the quotient nears the ends of
// the exponent range (d or 1/d subnormal, x/d subnormal, |x| < 2^-969, or
// exponent(x) - exponent(d) >= 768).  div_fixup decides the result from x
// and d alone when either is zero, infinite or NaN.  So for d in
// [2^-40, 2^94] and x = 0 or |x| in [2^-800, 2^340] (or x, d special), the
// steps without the scaling give the hardware quotient bit for bit, and one
// refined reciprocal serves every numerator over the same d: 4 instructions
// per quotient after 5 for the reciprocal, against 11 per quotient.
//
// The compiler's sqrt scales x < 2^-767 by 2^256 before its rsq-based
// Newton steps (and the result back) and passes +-0 and +Inf through; for x
// in [2^-767, DBL_MAX] both are no-ops, and sqrt_rng is the same steps
// without them: 10 instructions against 17.
//
// central_upwind_rng is central_upwind with these forms.  It is used only by
// a wave whose stage inputs (w, hu, hv of every cell the face stencils read)
// and bottom corners all lie in {0} U [2^-100, 2^40] in magnitude (checked as
// the rows arrive, in_band5, whose set is the subset {0} U [2^-100, 2^40)),
// with θ (linear reconstruction), ρ and g in
// [2^-10, 2^10] and depth_cutoff and κ in [2^-39, 2^10] (KParams::ranged).
// Then (exact differences of such values are multiples of 2^-152, rounded
// products keep their exponent sums, and cancellation in u +- c leaves at
// least an ulp of c >= sqrt(g*cutoff)):
//   face values:   0 or [2^-215, 2^51]; wet depths h > cutoff >= 2^-39
//   velocity quotient hu/h*:  x in {0} U [2^-215, 2^51], d = h* in [2^-40, 2^51]
//   c = sqrt(g h):  g h in [2^-49, 2^61]
//   a+, a-:         0 or [2^-78, 2^92];  da = a+ - a- in [κ, 2^93] unless degenerate
//   flux quotients: numerators 0 or [2^-712, 2^336], (a+ a-) in [2^-156, 2^184]
// which is inside the range above.  A degenerate face (|da| < κ) discards
// its quotients, as the reference's early return does.
struct Recip {
    double d, r;
};
__device__ __forceinline__ Recip recip_rng(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double t = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, t, r);
    t = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, t, r);
    return {d, r};
}
__device__ __forceinline__ double div_rng(double x, const Recip& R) {
    const double q = x * R.r;
    const double e = __builtin_fma(-R.d, q, x);
    return __builtin_amdgcn_div_fixup(__builtin_fma(e, R.r, q), R.d, x);
}
__device__ __forceinline__ double sqrt_rng(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
// |v| is 0 or in [2^-100, 2^40] (false for NaN, Inf and subnormals).
__device__ __forceinline__ bool in_band(double v) {
    const double a = __builtin_fabs(v);
    return a == 0.0 || (a >= 0x1p-100 && a <= 0x1p40);
}
// Every value 0 or in [2^-100, 2^40) in magnitude (a subset of in_band's
// set), tested through exponents: the hardware's frexp exponent e (v = m 2^e
// with |m| in [0.5, 1); 0 for 0, the true exponent for subnormals, 0 for Inf
// and NaN) lies in [-99, 40] exactly for 0 and the magnitudes in
// [2^-100, 2^40), and for Inf and NaN, which make the sum non-finite.  The
// exponents combine with integer min/max, so the test costs one lane mask
// instead of three per value.
__device__ __forceinline__ bool in_band5(double a, double b, double c, double d, double e) {
    const int ea = __builtin_amdgcn_frexp_exp(a), eb = __builtin_amdgcn_frexp_exp(b),
              ec = __builtin_amdgcn_frexp_exp(c), ed = __builtin_amdgcn_frexp_exp(d),
              ee = __builtin_amdgcn_frexp_exp(e);
    const int lo = min(min(min(ea, eb), min(ec, ed)), ee), hi = max(max(max(ea, eb), max(ec, ed)), ee);
    const double s = ((a + b) + (c + d)) + e;      // finite: no value is Inf or NaN (|v| < 2^40 cannot overflow it)
    return (unsigned)(lo + 99) <= 139u && (unsigned)(hi + 99) <= 139u && __builtin_isfinite(s);
}

// swe_2D.jl:43-81 with the ranged quotients and square root (h > cutoff).
template <class S>
__device__ __forceinline__ void phys_flux_rng(int d, S h, S hu, S hv, const PhysParams& p, double f[3],
                                              double e[2]) {
    const double hs = desingularize(h, p.kappa);
    const Recip R = recip_rng(hs);
    const double u = div_rng(hu, R);
    const double v = div_rng(hv, R);
    const S h2 = h * h;
    const double c = sqrt_rng(p.g * h);
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

// centralupwind.jl:48-77 as central_upwind, with the ranged forms.
template <class S>
__device__ __forceinline__ double central_upwind_rng(int d, const S m[3], const S q[3], const PhysParams& p,
                                                     double F[3]) {
    double fm[3] = {0.0, 0.0, 0.0}, em[2] = {0.0, 0.0};
    double fp[3] = {0.0, 0.0, 0.0}, ep[2] = {0.0, 0.0};
    if (m[0] > p.cutoff) phys_flux_rng(d, m[0], m[1], m[2], p, fm, em);
    if (q[0] > p.cutoff) phys_flux_rng(d, q[0], q[1], q[2], p, fp, ep);
    const double aplus = nz_max(nz_max(ep[0], em[0]), 0.0);
    const double aminus = nz_min(nz_min(ep[1], em[1]), 0.0);
    const double da = aplus - aminus;
    if (__builtin_fabs(da) < p.kappa) {
        F[0] = F[1] = F[2] = 0.0;
        return 0.0;
    }
    const Recip R = recip_rng(da);
    const double pm = div_rng(aplus * aminus, R);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const S diff = q[v] - m[v];
        F[v] = div_rng(aplus * fm[v] - aminus * fp[v], R) + pm * diff;
    }
    if (m[0] < p.cutoff && q[0] < p.cutoff) return 0.0;
    return nz_max(__builtin_fabs(aplus), __builtin_fabs(aminus));
}

// Both sides of a face wet with |h| >= κ: central_upwind_wet applies.
template <class S>
__device__ __forceinline__ bool face_is_wet(const S m[3], const S q[3], const PhysParams& p) {
    return m[0] > p.cutoff && q[0] > p.cutoff && (double)__builtin_fabs(m[0]) >= p.kappa &&
           (double)__builtin_fabs(q[0]) >= p.kappa;
}

// One face, with the straight-line evaluation when the whole wave's faces are
// wet (a wave-uniform branch) and the general evaluation otherwise.
template <int D, class S>
__device__ __forceinline__ double central_upwind_auto(const S m[3], const S q[3], const PhysParams& p, double F[3]) {
    if (__all(face_is_wet(m, q, p))) return central_upwind_wet<D>(m, q, p, F);
    return central_upwind(D, m, q, p, F);
}

}  // namespace sfv
