// swe_oracle.hpp — CPU ORACLE (test infrastructure, never the product path).
//
// A loop-faithful C++ restatement of the reference's shallow-water step
// (sintefmath/VolumeFluxes.jl, a.k.a. SinFVM).  It keeps the reference's pass
// structure one-to-one — zero → per direction {reconstruct, flux, bottom
// source} → non-directional sources → RK update → implicit friction →
// dry clamp → boundary condition — with the reference's buffers (left/right
// face buffers, wavespeed scalar field, n_substeps+1 state volumes) and its
// evaluation order, so that it is bit-reproducible against the reference in
// Float64.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it, to check results and as the timed CPU baseline.
//
// Storage type S is double (Float64 backend) or float (make_cpu_backend(Float32)).
// Values read from S-typed buffers stay S and parameters stay double, so C++'s
// float/double promotion reproduces Julia's Float32/Float64 promotion, and a
// cast to S at each store reproduces the reference's rounding on store.
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   jl_min/jl_max/minmod      test_minmod.jl:24-27 KATs, IEEE special cases
//   B lookups                 test_bottom_topography.jl KATs
//   update_bc (wall/periodic) test_update_bc.jl, test_periodic_bc_2d.jl index maps
//   reconstruct+flux+bottom   lake at rest to 1e-14 (test_lake_at_rest_2d.jl:97-98),
//                             2D(n,3) == 2D(3,n)^T == 1D bitwise (test_swashes.jl:317-328),
//                             mirror symmetry (test_shallow_water_dry_states_2d.jl:103-116),
//                             SWASHES dam-break closed forms (swashes.jl), mass conservation
//   rain / infiltration       test_rain.jl:56-73 KATs and 1.03 closed form,
//                             test_infiltration.jl:58-72
//   steppers                  convergence order 1 / 2 (test_timesteppers.jl:76-100)
//   friction (all 3 laws)     closed forms of the implicit update on a uniform periodic
//                             flow: FE and RK2 steps (stage input), CFL dt, speed 0, one
//                             flow direction, desingularised thin layer, 1D, to 1e-14
//                             relative (tests/test_oracle_friction.py); the cbrt of
//                             bsa2012/fcg2016 stays unpinned at ulp level (host libm)
//   Horton exp                parity unpinned at ulp level (exp of the host libm),
//                             f(0)=f0 and f(1e6)=fc pinned (test_infiltration.jl:60-61)
//   Neumann BC                restates the reference's right-side index map as written
//                             (bc.jl:53,58 copy ghost from ghost); pinned by index KATs only
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../include/sinfvm_hip.h"

namespace oracle {

// ---------------------------------------------------------------- Julia scalars
// Julia's max/min on IEEE floats (base/math.jl): NaN-propagating and
// -0.0 < +0.0.  Used everywhere the reference calls max/min.
inline double jl_max(double x, double y) {
    if (std::isnan(x)) return x;
    if (std::isnan(y)) return y;
    if (x > y) return x;
    if (y > x) return y;
    return std::signbit(x) ? y : x;
}
inline double jl_min(double x, double y) {
    if (std::isnan(x)) return x;
    if (std::isnan(y)) return y;
    if (x < y) return x;
    if (y < x) return y;
    return std::signbit(x) ? x : y;
}
inline float jl_max(float x, float y) {
    if (std::isnan(x)) return x;
    if (std::isnan(y)) return y;
    if (x > y) return x;
    if (y > x) return y;
    return std::signbit(x) ? y : x;
}
inline float jl_min(float x, float y) {
    if (std::isnan(x)) return x;
    if (std::isnan(y)) return y;
    if (x < y) return x;
    if (y < x) return y;
    return std::signbit(x) ? x : y;
}

// linearreconstruction.jl:28-35
inline double minmod(double a, double b, double c) {
    if ((a > 0) && (b > 0) && (c > 0)) return jl_min(jl_min(a, b), c);
    if (a < 0 && b < 0 && c < 0) return jl_max(jl_max(a, b), c);
    return 0.0;
}

// equation.jl:40-62, Brodtkorb & Holm (2021) eq. (26):
// h* = copysign(1,h) * max(|h|, min(h^2/(2κ) + κ/2, κ)).  `h^2` is h*h in h's
// own type (literal_pow), so a Float32 h squares in Float32.
template <class H>
inline double desingularize(H h, double kappa) {
    const H h2 = h * h;
    const double sgn = std::signbit(h) ? -1.0 : 1.0;   // copysign(1, h) is an Int
    return sgn * jl_max((double)std::fabs(h), jl_min(h2 / (2 * kappa) + kappa / 2.0, kappa));
}

struct Params {
    int nx, ny, gc, nxT, nyT;
    double dx, dy, g, rho, cutoff, kappa, theta, cfl, Cz;
    int bc, rec, stepper, bottom, friction, rain, infil, rain_first, embedded_1d;
    double rain_rate, infil_rate, f0, fc, hk, factor_scalar;
    std::vector<double> rain_rates, rain_times, factor;
    std::vector<double> rain_field;   // FunctionalRain: rain_function(t, cell_center) of the current step
};

inline Params make_params(const sfv_config& c) {
    Params p;
    p.nx = c.nx; p.ny = c.ny; p.gc = c.gc;
    p.nxT = c.nx + 2 * c.gc; p.nyT = c.ny + 2 * c.gc;
    p.dx = c.dx; p.dy = c.dy; p.g = c.g; p.rho = c.rho;
    p.cutoff = c.depth_cutoff; p.kappa = c.desingularizing_kappa;
    p.theta = c.theta; p.cfl = c.cfl; p.Cz = c.Cz;
    p.bc = c.bc; p.rec = c.reconstruction; p.stepper = c.stepper;
    p.bottom = c.bottom_source; p.friction = c.friction; p.rain = c.rain;
    p.infil = c.infiltration; p.rain_first = c.rain_before_infiltration;
    p.rain_rate = c.rain_rate; p.infil_rate = c.infiltration_rate;
    p.f0 = c.horton_f0; p.fc = c.horton_fc; p.hk = c.horton_k;
    p.factor_scalar = c.infiltration_factor_scalar;
    p.embedded_1d = c.embedded_1d;
    if (c.rain == SFV_RAIN_TIME_TABLE && c.rain_table_len > 0) {
        p.rain_rates.assign(c.rain_table_rates, c.rain_table_rates + c.rain_table_len);
        p.rain_times.assign(c.rain_table_times, c.rain_table_times + c.rain_table_len);
    }
    if (c.rain == SFV_RAIN_FIELD) p.rain_field.assign((size_t)p.nxT * p.nyT, 0.0);
    if (c.infiltration == SFV_INFILTRATION_HORTON) {
        const size_t n = (size_t)p.nxT * p.nyT;
        if (c.infiltration_factor) p.factor.assign(c.infiltration_factor, c.infiltration_factor + n);
        else p.factor.assign(n, c.infiltration_factor_scalar);
    }
    return p;
}

// source_term_rain.jl:30 (ConstantRain) and :45-54 (TimeDependentRain)
inline double compute_rain(const Params& p, double t) {
    if (p.rain == SFV_RAIN_CONSTANT) return p.rain_rate / 3600;
    if (p.rain == SFV_RAIN_TIME_TABLE) {
        size_t index = 0;                          // 0-based restatement of the 1-based scan
        while (index + 1 < p.rain_times.size()) {
            if (t < p.rain_times[index + 1]) break;
            index += 1;
        }
        return p.rain_rates[index] / 3600;
    }
    return 0.0;
}
// source_term_infiltration.jl:68-81
inline double compute_infiltration(const Params& p, double t, size_t idx) {
    if (p.infil == SFV_INFILTRATION_HORTON)
        return p.factor[idx] * (p.fc + (p.f0 - p.fc) * std::exp(-p.hk * t));
    if (p.infil == SFV_INFILTRATION_CONSTANT) return p.infil_rate;
    return 0.0;
}

// friction.jl:30-43
inline double friction_law(int law, double c, double h, double speed) {
    if (law == SFV_FRICTION_BSA2012) { double denom = std::cbrt(h) * h; return -c * speed / denom; }
    if (law == SFV_FRICTION_FCG2016) { double denom = std::cbrt(h) * h * h; return -c * speed / denom; }
    double denom = h * h;                                   // friction_bh2021
    return -c * speed / denom;
}

// ============================================================== 2D simulator
// Mirrors Simulator + ConservedSystem for ShallowWaterEquations on a
// CartesianGrid{2} (simulator.jl:60-91, system.jl:24-54).
template <class S>
struct Sim2D {
    Params p;
    size_t ncell;                       // nxT * nyT
    std::vector<double> B;              // corners, (nxT+1) x (nyT+1)
    std::vector<std::vector<S>> out;    // substep_outputs: n_substeps+1 volumes of 3 planes
    std::vector<S> left, right;         // system.left_buffer / right_buffer
    std::vector<S> ws;                  // system.wavespeeds
    S current_timestep = 0;

    explicit Sim2D(const sfv_config& c) : p(make_params(c)) {
        ncell = (size_t)p.nxT * p.nyT;
        B.assign((size_t)(p.nxT + 1) * (p.nyT + 1), 0.0);
        const int nsub = (p.stepper == SFV_RUNGE_KUTTA2) ? 2 : 1;
        out.assign(nsub + 1, std::vector<S>(3 * ncell, S(0)));
        left.assign(3 * ncell, S(0));
        right.assign(3 * ncell, S(0));
        ws.assign(ncell, S(0));
    }
    inline size_t id(int i, int j) const { return (size_t)i + (size_t)p.nxT * j; }
    inline double Bc(int i, int j) const { return B[(size_t)i + (size_t)(p.nxT + 1) * j]; }

    // bottom_topography.jl:97-101 (0-based corner (i,j) is the reference's B[i+1, j+1])
    inline double B_cell(int i, int j) const {
        // bottom_topography.jl:97; the 1D lookup (:92) for a 1D run carried by an (nx, 3) grid
        if (p.embedded_1d) return 0.5 * (Bc(i, j) + Bc(i + 1, j));
        return 0.25 * (Bc(i, j) + Bc(i + 1, j) + Bc(i, j + 1) + Bc(i + 1, j + 1));
    }
    inline double B_face_left(int i, int j, int d) const {
        return d == 0 ? 0.5 * (Bc(i, j) + Bc(i, j + 1)) : 0.5 * (Bc(i, j) + Bc(i + 1, j));
    }
    inline double B_face_right(int i, int j, int d) const {
        return d == 0 ? 0.5 * (Bc(i + 1, j) + Bc(i + 1, j + 1)) : 0.5 * (Bc(i, j + 1) + Bc(i + 1, j + 1));
    }

    // ---- reconstruction, linearreconstruction.jl:58-102 (SWE variant) and
    //      noreconstruction.jl:30-46; iteration domain for_each_inner_cell
    //      with ghostcells=1 (kernel_abstractions.jl:66-75, grid.jl:213-244).
    void reconstruct(const std::vector<S>& in, int d) {
        const size_t n = ncell;
        if (p.rec == SFV_NO_RECONSTRUCTION) {
            #pragma omp parallel for schedule(static)
            for (int j = 0; j < p.nyT; ++j)
                for (int i = 0; i < p.nxT; ++i) {
                    size_t k = id(i, j);
                    for (int v = 0; v < 3; ++v) { left[v * n + k] = in[v * n + k]; right[v * n + k] = in[v * n + k]; }
                }
            #pragma omp parallel for schedule(static)
            for (int j = 0; j < p.nyT; ++j)
                for (int i = 0; i < p.nxT; ++i) {
                    size_t k = id(i, j);
                    left[k] = (S)(in[k] - B_cell(i, j));
                    right[k] = (S)(in[k] - B_cell(i, j));
                }
            return;
        }
        const int i_lo = (d == 0) ? 1 : p.gc, i_hi = (d == 0) ? p.nxT - 1 : p.nxT - p.gc;
        const int j_lo = (d == 1) ? 1 : p.gc, j_hi = (d == 1) ? p.nyT - 1 : p.nyT - p.gc;
        const size_t step = (d == 0) ? 1 : (size_t)p.nxT;
        #pragma omp parallel for schedule(static)
        for (int j = j_lo; j < j_hi; ++j)
            for (int i = i_lo; i < i_hi; ++i) {
                const size_t km = id(i, j), kl = km - step, kr = km + step;
                double slope[3];
                for (int v = 0; v < 3; ++v) {
                    // minmod_slope (linearreconstruction.jl:37-43); differences in S
                    const S fwd = in[v * n + kr] - in[v * n + km];
                    const S bwd = in[v * n + km] - in[v * n + kl];
                    const double central = (fwd + bwd) / 2.0;
                    slope[v] = minmod(p.theta * fwd, central, p.theta * bwd);
                }
                const double Bl = B_face_left(i, j, d), Br = B_face_right(i, j, d);
                const S w = in[km];
                if (w - 0.5 * slope[0] < Bl) slope[0] = 2.0 * (w - Bl);
                else if (w + 0.5 * slope[0] < Br) slope[0] = 2.0 * (Br - w);
                for (int v = 0; v < 3; ++v) {
                    left[v * n + km] = (S)(in[v * n + km] - 0.5 * slope[v]);
                    right[v * n + km] = (S)(in[v * n + km] + 0.5 * slope[v]);
                }
                left[km] = (S)(left[km] - Bl);
                right[km] = (S)(right[km] - Br);
            }
    }

    // swe_2D.jl:43-65 physical flux and :69-81 eigenvalues, for a face state
    // (h, hu, hv) of storage type S.  Returns flux f[3] and eigenvalues e[0..1].
    inline void phys(int d, S h, S hu, S hv, double f[3], double e[2]) const {
        const double hs = desingularize(h, p.kappa);
        const double u = hu / hs, v = hv / hs;
        const S h2 = h * h;
        if (d == 0) {
            f[0] = p.rho * h * u;
            f[1] = p.rho * h * u * u + 0.5 * p.rho * p.g * h2;
            f[2] = p.rho * h * u * v;
            const double c = std::sqrt(p.g * h);
            e[0] = u + c; e[1] = u - c;
        } else {
            f[0] = p.rho * h * v;
            f[1] = p.rho * h * u * v;
            f[2] = p.rho * h * v * v + 0.5 * p.rho * p.g * h2;
            const double c = std::sqrt(p.g * h);
            e[0] = v + c; e[1] = v - c;
        }
    }

    // centralupwind.jl:48-77, practical SWE variant.  m = faceminus, p = faceplus.
    inline double central_upwind(int d, const S m[3], const S q[3], double F[3]) const {
        double fm[3] = {0, 0, 0}, em[2] = {0, 0}, fp[3] = {0, 0, 0}, ep[2] = {0, 0};
        if (m[0] > p.cutoff) phys(d, m[0], m[1], m[2], fm, em);
        if (q[0] > p.cutoff) phys(d, q[0], q[1], q[2], fp, ep);
        const double aplus = jl_max(jl_max(ep[0], em[0]), 0.0);
        const double aminus = jl_min(jl_min(ep[1], em[1]), 0.0);
        if (std::fabs(aplus - aminus) < p.kappa) { F[0] = F[1] = F[2] = 0.0; return 0.0; }
        for (int v = 0; v < 3; ++v) {
            const S diff = q[v] - m[v];
            F[v] = (aplus * fm[v] - aminus * fp[v]) / (aplus - aminus) + ((aplus * aminus) / (aplus - aminus)) * diff;
        }
        if (m[0] < p.cutoff && q[0] < p.cutoff) return 0.0;
        return jl_max(std::fabs(aplus), std::fabs(aminus));
    }

    // numericalflux.jl:21-33 over for_each_inner_cell (interior cells);
    // returns maximum(wavespeeds) over the whole buffer.
    S compute_flux(std::vector<S>& o, int d) {
        const size_t n = ncell;
        const double dxd = (d == 0) ? p.dx : p.dy;
        const double inv = 1 / dxd;
        const size_t step = (d == 0) ? 1 : (size_t)p.nxT;
        #pragma omp parallel for schedule(static)
        for (int j = p.gc; j < p.nyT - p.gc; ++j)
            for (int i = p.gc; i < p.nxT - p.gc; ++i) {
                const size_t km = id(i, j), kl = km - step, kr = km + step;
                S Rm[3], Lr[3], Rl[3], Lm[3];
                for (int v = 0; v < 3; ++v) {
                    Rm[v] = right[v * n + km]; Lr[v] = left[v * n + kr];
                    Rl[v] = right[v * n + kl]; Lm[v] = left[v * n + km];
                }
                double Fr[3], Fl[3];
                const double sr = central_upwind(d, Rm, Lr, Fr);
                const double sl = central_upwind(d, Rl, Lm, Fl);
                for (int v = 0; v < 3; ++v) o[v * n + km] = (S)(o[v * n + km] - inv * (Fr[v] - Fl[v]));
                ws[km] = (S)jl_max(sr, sl);
            }
        // maximum(wavespeeds): max is exact, so a per-row reduction is order-free.
        std::vector<S> rowmax(p.nyT, S(0));
        #pragma omp parallel for schedule(static)
        for (int j = 0; j < p.nyT; ++j) {
            S m = ws[id(0, j)];
            for (int i = 1; i < p.nxT; ++i) m = jl_max(m, ws[id(i, j)]);
            rowmax[j] = m;
        }
        S m = rowmax[0];
        for (int j = 1; j < p.nyT; ++j) m = jl_max(m, rowmax[j]);
        return m;
    }

    // source_term_bottom.jl:24-41
    void bottom_source(std::vector<S>& o, int d) {
        const size_t n = ncell;
        const double dxd = (d == 0) ? p.dx : p.dy;
        const int v = 1 + d;
        #pragma omp parallel for schedule(static)
        for (int j = p.gc; j < p.nyT - p.gc; ++j)
            for (int i = p.gc; i < p.nxT - p.gc; ++i) {
                const size_t k = id(i, j);
                const double Br = B_face_right(i, j, d), Bl = B_face_left(i, j, d);
                const S hsum = right[k] + left[k];
                o[v * n + k] = (S)(o[v * n + k] + -p.g * ((Br - Bl) / dxd) * (hsum / 2.0));
            }
    }

    // source_term_rain.jl:71-76 and source_term_infiltration.jl:83-88 (all cells)
    void nondirectional_sources(std::vector<S>& o, double t) {
        auto rain = [&]() {
            if (p.rain == SFV_RAIN_NONE) return;
            if (p.rain == SFV_RAIN_FIELD) {   // source_term_rain.jl:64-68, values set by the caller for t
                #pragma omp parallel for schedule(static)
                for (long k = 0; k < (long)ncell; ++k) o[k] = (S)(o[k] + p.rain_field[k]);
                return;
            }
            const double r = compute_rain(p, t);
            #pragma omp parallel for schedule(static)
            for (long k = 0; k < (long)ncell; ++k) o[k] = (S)(o[k] + r);
        };
        auto infil = [&]() {
            if (p.infil == SFV_INFILTRATION_NONE) return;
            #pragma omp parallel for schedule(static)
            for (long k = 0; k < (long)ncell; ++k) o[k] = (S)(o[k] - compute_infiltration(p, t, (size_t)k));
        };
        if (p.rain_first) { rain(); infil(); } else { infil(); rain(); }
    }

    // friction.jl:62-84 (all cells; state is the substep input)
    void implicit_friction(std::vector<S>& o, const std::vector<S>& st, S dt) {
        if (p.friction == SFV_FRICTION_NONE) return;
        const size_t n = ncell;
        #pragma omp parallel for schedule(static)
        for (int j = 0; j < p.nyT; ++j)
            for (int i = 0; i < p.nxT; ++i) {
                const size_t k = id(i, j);
                const double Bm = B_cell(i, j);
                const double hs = desingularize(st[k] - Bm, p.kappa);
                const double u = st[n + k] / hs, v = st[2 * n + k] / hs;
                double speed = std::sqrt(u * u + v * v);
                if (speed == 0.0) speed = 0.0;
                const double fs = friction_law(p.friction, p.Cz, hs, speed);
                o[k] = (S)(o[k] / (1 - dt * 0.0));
                o[n + k] = (S)(o[n + k] / (1 - dt * fs));
                o[2 * n + k] = (S)(o[2 * n + k] / (1 - dt * fs));
            }
    }

    // timestepper.jl:53-63 (all cells)
    void post_proc(std::vector<S>& o) {
        const size_t n = ncell;
        #pragma omp parallel for schedule(static)
        for (int j = 0; j < p.nyT; ++j)
            for (int i = 0; i < p.nxT; ++i) {
                const size_t k = id(i, j);
                const double b = B_cell(i, j);
                if (o[k] - b < p.cutoff) {
                    o[k] = (S)jl_max((double)o[k], b);
                    o[n + k] = (S)0.0;
                    o[2 * n + k] = (S)0.0;
                }
            }
    }

    // bc.jl:29-110 for CartesianGrid{2}: X ghosts (interior rows) then Y ghosts
    // (interior columns); corners are never visited (for_each_ghost_cell,
    // kernel_abstractions.jl:77-85, ghost_cells grid.jl:246-253).
    void update_bc(std::vector<S>& U) {
        const size_t n = ncell;
        const int gc = p.gc, nxT = p.nxT, nyT = p.nyT;
        auto put = [&](size_t dst, size_t src, int neg) {
            for (int v = 0; v < 3; ++v) {
                const S val = U[v * n + src];
                U[v * n + dst] = (v == neg) ? (S)(-val) : val;
            }
        };
        for (int j = gc; j < nyT - gc; ++j)
            for (int g = 1; g <= gc; ++g) {       // 1-based ghost index as in bc.jl
                if (p.bc == SFV_BC_WALL) {
                    put(id(g - 1, j), id(2 * gc - g, j), 1);
                    put(id(nxT - gc + g - 1, j), id(nxT - gc - g, j), 1);
                } else if (p.bc == SFV_BC_PERIODIC) {
                    put(id(g - 1, j), id(nxT + g - 2 * gc - 1, j), -1);
                    put(id(nxT - (gc - g) - 1, j), id(gc + g - 1, j), -1);
                } else {                              // NeumannBC, index map exactly as bc.jl:52-53
                    put(id(g - 1, j), id(2 * gc - g, j), -1);
                    put(id(nxT - (gc - g) - 1, j), id(nxT - g, j), -1);
                }
            }
        for (int i = gc; i < nxT - gc; ++i)
            for (int g = 1; g <= gc; ++g) {
                if (p.bc == SFV_BC_WALL) {
                    put(id(i, g - 1), id(i, 2 * gc - g), 2);
                    put(id(i, nyT - gc + g - 1), id(i, nyT - gc - g), 2);
                } else if (p.bc == SFV_BC_PERIODIC) {
                    put(id(i, g - 1), id(i, nyT + g - 2 * gc - 1), -1);
                    put(id(i, nyT - (gc - g) - 1), id(i, gc + g - 1), -1);
                } else {                              // bc.jl:57-58
                    put(id(i, g - 1), id(i, 2 * gc - g), -1);
                    put(id(i, nyT - (gc - g) - 1), id(i, nyT - g), -1);
                }
            }
    }

    // rungekutta.jl:26-52 / forwardeuler.jl:28-47.  Returns the dt used for
    // the RK update: in stage 1 the Float64 value of timestep_computer, in
    // later stages the stored current_timestep (realtype), so that a Float32
    // backend does stage-2 `in + dt*out` in Float32 arithmetic.
    double do_substep(int s, S dt_stored, double max_dt, double t) {
        std::vector<S>& o = out[s + 1];
        const std::vector<S>& in = out[s];
        std::fill(o.begin(), o.end(), S(0));
        // add_time_derivative! (system.jl:62-75)
        S wave[2];
        for (int d = 0; d < 2; ++d) {
            reconstruct(in, d);
            wave[d] = compute_flux(o, d);
            if (p.bottom) bottom_source(o, d);
        }
        nondirectional_sources(o, t);
        const size_t n3 = 3 * ncell;
        double dt = (double)dt_stored;
        if (s == 0) {
            // timestep_computer (simulator.jl:147-150); cfl is stored as realtype
            const S cfl = (S)p.cfl;
            const double a = p.dx / wave[0], b = p.dy / wave[1];
            dt = jl_min(cfl * jl_min(a, b), max_dt);
            #pragma omp parallel for schedule(static)
            for (long k = 0; k < (long)n3; ++k) o[k] = (S)(in[k] + dt * o[k]);
        } else {
            #pragma omp parallel for schedule(static)
            for (long k = 0; k < (long)n3; ++k) o[k] = (S)(in[k] + dt_stored * o[k]);
        }
        if (p.stepper == SFV_RUNGE_KUTTA2 && s == 1) {
            const std::vector<S>& first = out[0];
            #pragma omp parallel for schedule(static)
            for (long k = 0; k < (long)n3; ++k) o[k] = (S)(0.5 * (first[k] + o[k]));
        }
        return dt;
    }

    // perform_step! (simulator.jl:145-177).  Returns current_timestep.
    double perform_step(double max_dt, double t) {
        const int nsub = (p.stepper == SFV_RUNGE_KUTTA2) ? 2 : 1;
        for (int s = 0; s < nsub; ++s) {
            const double dt = do_substep(s, current_timestep, max_dt, t);
            current_timestep = (S)dt;
            implicit_friction(out[s + 1], out[s], current_timestep);
            post_proc(out[s + 1]);
            update_bc(out[s + 1]);
        }
        std::swap(out[0], out[nsub]);
        return (double)current_timestep;
    }

    // set_current_state! (simulator.jl:119-135): interior then BC.
    void set_interior(const double* U, int nxi, int nyi) {
        const size_t n = ncell;
        for (int v = 0; v < 3; ++v)
            for (int j = 0; j < nyi; ++j)
                for (int i = 0; i < nxi; ++i)
                    out[0][v * n + id(i + p.gc, j + p.gc)] = (S)U[(size_t)v * nxi * nyi + (size_t)j * nxi + i];
        update_bc(out[0]);
    }
};

// ============================================================== 1D simulator
// ShallowWaterEquations1D on CartesianGrid{1}: the same scheme with state
// (w, hu) and B of length nxT+1 (swe_1D.jl, bc.jl:63-78, timestepper.jl:29-39,
// friction.jl:45-60, bottom_topography.jl:92-94).  CPU oracle only.
struct Sim1D {
    Params p;
    int nxT;
    std::vector<double> B;
    std::vector<std::vector<double>> out;
    std::vector<double> left, right, ws;
    double current_timestep = 0;

    explicit Sim1D(const sfv_config& c) : p(make_params(c)) {
        nxT = p.nxT;
        B.assign(nxT + 1, 0.0);
        const int nsub = (p.stepper == SFV_RUNGE_KUTTA2) ? 2 : 1;
        out.assign(nsub + 1, std::vector<double>(2 * nxT, 0.0));
        left.assign(2 * nxT, 0.0); right.assign(2 * nxT, 0.0); ws.assign(nxT, 0.0);
    }
    double B_cell(int i) const { return 0.5 * (B[i] + B[i + 1]); }

    void reconstruct(const std::vector<double>& in) {
        const int n = nxT;
        if (p.rec == SFV_NO_RECONSTRUCTION) {
            for (int i = 0; i < n; ++i) for (int v = 0; v < 2; ++v) { left[v * n + i] = right[v * n + i] = in[v * n + i]; }
            for (int i = 0; i < n; ++i) { left[i] = in[i] - B_cell(i); right[i] = in[i] - B_cell(i); }
            return;
        }
        for (int i = 1; i < n - 1; ++i) {
            double slope[2];
            for (int v = 0; v < 2; ++v) {
                const double fwd = in[v * n + i + 1] - in[v * n + i];
                const double bwd = in[v * n + i] - in[v * n + i - 1];
                slope[v] = minmod(p.theta * fwd, (fwd + bwd) / 2.0, p.theta * bwd);
            }
            const double Bl = B[i], Br = B[i + 1], w = in[i];
            if (w - 0.5 * slope[0] < Bl) slope[0] = 2.0 * (w - Bl);
            else if (w + 0.5 * slope[0] < Br) slope[0] = 2.0 * (Br - w);
            for (int v = 0; v < 2; ++v) {
                left[v * n + i] = in[v * n + i] - 0.5 * slope[v];
                right[v * n + i] = in[v * n + i] + 0.5 * slope[v];
            }
            left[i] -= Bl; right[i] -= Br;
        }
    }
    double cu(const double m[2], const double q[2], double F[2]) const {
        double fm[2] = {0, 0}, em[2] = {0, 0}, fp[2] = {0, 0}, ep[2] = {0, 0};
        auto phys = [&](double h, double hu, double f[2], double e[2]) {
            const double u = hu / desingularize(h, p.kappa);
            f[0] = p.rho * h * u;
            f[1] = p.rho * h * u * u + 0.5 * p.rho * p.g * (h * h);
            const double c = std::sqrt(p.g * h);
            e[0] = u + c; e[1] = u - c;
        };
        if (m[0] > p.cutoff) phys(m[0], m[1], fm, em);
        if (q[0] > p.cutoff) phys(q[0], q[1], fp, ep);
        const double aplus = jl_max(jl_max(ep[0], em[0]), 0.0);
        const double aminus = jl_min(jl_min(ep[1], em[1]), 0.0);
        if (std::fabs(aplus - aminus) < p.kappa) { F[0] = F[1] = 0.0; return 0.0; }
        for (int v = 0; v < 2; ++v)
            F[v] = (aplus * fm[v] - aminus * fp[v]) / (aplus - aminus) + ((aplus * aminus) / (aplus - aminus)) * (q[v] - m[v]);
        if (m[0] < p.cutoff && q[0] < p.cutoff) return 0.0;
        return jl_max(std::fabs(aplus), std::fabs(aminus));
    }
    double compute_flux(std::vector<double>& o) {
        const int n = nxT;
        const double inv = 1 / p.dx;
        for (int i = p.gc; i < n - p.gc; ++i) {
            double Rm[2] = {right[i], right[n + i]}, Lr[2] = {left[i + 1], left[n + i + 1]};
            double Rl[2] = {right[i - 1], right[n + i - 1]}, Lm[2] = {left[i], left[n + i]};
            double Fr[2], Fl[2];
            const double sr = cu(Rm, Lr, Fr), sl = cu(Rl, Lm, Fl);
            for (int v = 0; v < 2; ++v) o[v * n + i] = o[v * n + i] - inv * (Fr[v] - Fl[v]);
            ws[i] = jl_max(sr, sl);
        }
        double m = ws[0];
        for (int i = 1; i < n; ++i) m = jl_max(m, ws[i]);
        return m;
    }
    void update_bc(std::vector<double>& U) {
        const int n = nxT, gc = p.gc;
        for (int g = 1; g <= gc; ++g) {
            if (p.bc == SFV_BC_WALL) {
                int lg = g - 1, li = 2 * gc - g, rg = n - gc + g - 1, ri = n - gc - g;
                U[lg] = U[li]; U[n + lg] = -U[n + li];
                U[rg] = U[ri]; U[n + rg] = -U[n + ri];
            } else if (p.bc == SFV_BC_PERIODIC) {
                int lg = g - 1, li = n + g - 2 * gc - 1, rg = n - (gc - g) - 1, ri = gc + g - 1;
                U[lg] = U[li]; U[n + lg] = U[n + li];
                U[rg] = U[ri]; U[n + rg] = U[n + ri];
            } else {
                int lg = g - 1, li = 2 * gc - g, rg = n - (gc - g) - 1, ri = n - g;
                U[lg] = U[li]; U[n + lg] = U[n + li];
                U[rg] = U[ri]; U[n + rg] = U[n + ri];
            }
        }
    }
    double perform_step(double max_dt, double t) {
        const int n = nxT;
        const int nsub = (p.stepper == SFV_RUNGE_KUTTA2) ? 2 : 1;
        for (int s = 0; s < nsub; ++s) {
            std::vector<double>& o = out[s + 1];
            const std::vector<double>& in = out[s];
            std::fill(o.begin(), o.end(), 0.0);
            reconstruct(in);
            const double wave = compute_flux(o);
            if (p.bottom) {
                for (int i = p.gc; i < n - p.gc; ++i)
                    o[n + i] = o[n + i] + -p.g * ((B[i + 1] - B[i]) / p.dx) * ((right[i] + left[i]) / 2.0);
            }
            const double r = compute_rain(p, t);
            auto rain = [&]() { if (p.rain) for (int i = 0; i < n; ++i) o[i] = o[i] + r; };
            auto infil = [&]() { if (p.infil) for (int i = 0; i < n; ++i) o[i] = o[i] - compute_infiltration(p, t, (size_t)i); };
            if (p.rain_first) { rain(); infil(); } else { infil(); rain(); }
            double dt = current_timestep;
            if (s == 0) dt = jl_min(p.cfl * (p.dx / wave), max_dt);
            for (int k = 0; k < 2 * n; ++k) o[k] = in[k] + dt * o[k];
            if (nsub == 2 && s == 1)
                for (int k = 0; k < 2 * n; ++k) o[k] = 0.5 * (out[0][k] + o[k]);
            current_timestep = dt;
            if (p.friction) {
                for (int i = 0; i < n; ++i) {
                    const double hs = desingularize(in[i] - B_cell(i), p.kappa);
                    const double u = in[n + i] / hs;
                    double speed = std::sqrt(u * u);
                    if (speed == 0.0) speed = 0.0;
                    const double fs = friction_law(p.friction, p.Cz, hs, speed);
                    o[i] = o[i] / (1 - dt * 0.0);
                    o[n + i] = o[n + i] / (1 - dt * fs);
                }
            }
            for (int i = 0; i < n; ++i) {
                const double b = B_cell(i);
                if (o[i] - b < p.cutoff) { o[i] = jl_max(o[i], b); o[n + i] = 0.0; }
            }
            update_bc(o);
        }
        std::swap(out[0], out[nsub]);
        return current_timestep;
    }
};

}  // namespace oracle
