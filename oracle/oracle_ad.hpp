// oracle_ad.hpp — CPU ORACLE (test infrastructure, never the product path):
// forward-mode differentiation of the reference's 2D shallow-water step.
//
// A compact restatement of the same step as swe_oracle.hpp's Sim2D (Float64
// only: reconstruction, central-upwind flux, bottom source, rain and
// infiltration, RK2/FE update, implicit friction, dry clamp, boundary
// condition; simulator.jl:145-177), written once on a generic number type T.
// With T = double it must equal Sim2D bit for bit (pinned by a test); with
// T = Dual<N> it carries N directional derivatives through every step, the
// way ForwardDiff carries Duals through the reference (branches decided on
// values, derivatives of the taken branch, dt differentiated through the
// maximum wave speed).  Used to check the GPU adjoint: <dJ/dp, δp> from the
// adjoint must equal the directional derivative computed here.
#pragma once

#include <cmath>
#include <vector>

#include "../include/sinfvm_hip.h"

namespace oracle_ad {

template <int N>
struct Dual {
    double v = 0.0;
    double d[N] = {};
    Dual() = default;
    Dual(double x) : v(x) {}
};
inline double val(double x) { return x; }
template <int N> double val(const Dual<N>& x) { return x.v; }

#define OAD_BIN(op, expr_d)                                                              \
    template <int N> Dual<N> operator op(const Dual<N>& a, const Dual<N>& b) {            \
        Dual<N> r(a.v op b.v);                                                           \
        for (int k = 0; k < N; ++k) r.d[k] = expr_d;                                     \
        return r;                                                                        \
    }                                                                                    \
    template <int N> Dual<N> operator op(const Dual<N>& a, double b) { return a op Dual<N>(b); } \
    template <int N> Dual<N> operator op(double a, const Dual<N>& b) { return Dual<N>(a) op b; }
OAD_BIN(+, a.d[k] + b.d[k])
OAD_BIN(-, a.d[k] - b.d[k])
OAD_BIN(*, a.d[k] * b.v + a.v * b.d[k])
OAD_BIN(/, (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v))
#undef OAD_BIN
template <int N> Dual<N> operator-(const Dual<N>& a) { return Dual<N>(0.0) - a; }
template <int N> Dual<N> sqrt_(const Dual<N>& a) {
    Dual<N> r(std::sqrt(a.v));
    for (int k = 0; k < N; ++k) r.d[k] = r.v > 0 ? a.d[k] / (2.0 * r.v) : 0.0;
    return r;
}
inline double sqrt_(double a) { return std::sqrt(a); }
template <int N> Dual<N> cbrt_(const Dual<N>& a) {
    Dual<N> r(std::cbrt(a.v));
    for (int k = 0; k < N; ++k) r.d[k] = a.v != 0 ? a.d[k] * r.v / (3.0 * a.v) : 0.0;
    return r;
}
inline double cbrt_(double a) { return std::cbrt(a); }
template <class T> T abs_(const T& a) { return std::signbit(val(a)) ? -a : a; }
inline double abs_(double a) { return std::fabs(a); }

// Julia max/min (base/math.jl), choosing one argument whole.
template <class T> T jmax(const T& x, const T& y) {
    const double a = val(x), b = val(y);
    if (std::isnan(a)) return x;
    if (std::isnan(b)) return y;
    if (a > b) return x;
    if (b > a) return y;
    return std::signbit(a) ? y : x;
}
template <class T> T jmin(const T& x, const T& y) {
    const double a = val(x), b = val(y);
    if (std::isnan(a)) return x;
    if (std::isnan(b)) return y;
    if (a < b) return x;
    if (b < a) return y;
    return std::signbit(a) ? x : y;
}

template <class T>
struct Step2D {
    sfv_config c;
    int nxT, nyT;
    size_t n;
    std::vector<T> B;                    // corners
    std::vector<T> U[3], L, R, ws;       // substep outputs 0..2, face buffers, wavespeeds
    std::vector<double> rain_rates, rain_times, factor;
    T Cz;
    T dt_cur = T(0.0);

    explicit Step2D(const sfv_config& cfg) : c(cfg) {
        nxT = c.nx + 2 * c.gc;
        nyT = c.ny + 2 * c.gc;
        n = (size_t)nxT * nyT;
        B.assign((size_t)(nxT + 1) * (nyT + 1), T(0.0));
        for (auto& u : U) u.assign(3 * n, T(0.0));
        L.assign(3 * n, T(0.0));
        R.assign(3 * n, T(0.0));
        ws.assign(n, T(0.0));
        Cz = T(c.Cz);
        if (c.rain == SFV_RAIN_TIME_TABLE) {
            rain_rates.assign(c.rain_table_rates, c.rain_table_rates + c.rain_table_len);
            rain_times.assign(c.rain_table_times, c.rain_table_times + c.rain_table_len);
        }
        if (c.infiltration == SFV_INFILTRATION_HORTON) {
            if (c.infiltration_factor) factor.assign(c.infiltration_factor, c.infiltration_factor + n);
            else factor.assign(n, c.infiltration_factor_scalar);
        }
    }
    size_t id(int i, int j) const { return (size_t)i + (size_t)nxT * j; }
    const T& Bc(int i, int j) const { return B[(size_t)i + (size_t)(nxT + 1) * j]; }
    // bottom_topography.jl:97-101
    T B_cell(int i, int j) const { return 0.25 * (Bc(i, j) + Bc(i + 1, j) + Bc(i, j + 1) + Bc(i + 1, j + 1)); }
    T B_left(int i, int j, int d) const { return d == 0 ? 0.5 * (Bc(i, j) + Bc(i, j + 1)) : 0.5 * (Bc(i, j) + Bc(i + 1, j)); }
    T B_right(int i, int j, int d) const {
        return d == 0 ? 0.5 * (Bc(i + 1, j) + Bc(i + 1, j + 1)) : 0.5 * (Bc(i, j + 1) + Bc(i + 1, j + 1));
    }

    // equation.jl:56
    T desing(const T& h) const {
        const double k = c.desingularizing_kappa;
        return std::copysign(1.0, val(h)) * jmax(abs_(h), jmin(h * h / (2 * k) + k / 2.0, T(k)));
    }
    // linearreconstruction.jl:28-35
    static T minmod(const T& a, const T& b, const T& cc) {
        if (val(a) > 0 && val(b) > 0 && val(cc) > 0) return jmin(jmin(a, b), cc);
        if (val(a) < 0 && val(b) < 0 && val(cc) < 0) return jmax(jmax(a, b), cc);
        return T(0.0);
    }
    void reconstruct(const std::vector<T>& in, int d) {
        const int i_lo = d == 0 ? 1 : c.gc, i_hi = d == 0 ? nxT - 1 : nxT - c.gc;
        const int j_lo = d == 1 ? 1 : c.gc, j_hi = d == 1 ? nyT - 1 : nyT - c.gc;
        const size_t step = d == 0 ? 1 : (size_t)nxT;
        #pragma omp parallel for schedule(static)
        for (int j = j_lo; j < j_hi; ++j)
            for (int i = i_lo; i < i_hi; ++i) {
                const size_t km = id(i, j), kl = km - step, kr = km + step;
                T slope[3];
                for (int v = 0; v < 3; ++v) {
                    const T fwd = in[v * n + kr] - in[v * n + km];
                    const T bwd = in[v * n + km] - in[v * n + kl];
                    slope[v] = minmod(c.theta * fwd, (fwd + bwd) / 2.0, c.theta * bwd);
                }
                const T Bl = B_left(i, j, d), Br = B_right(i, j, d);
                const T w = in[km];
                if (val(w - 0.5 * slope[0]) < val(Bl)) slope[0] = 2.0 * (w - Bl);
                else if (val(w + 0.5 * slope[0]) < val(Br)) slope[0] = 2.0 * (Br - w);
                for (int v = 0; v < 3; ++v) {
                    L[v * n + km] = in[v * n + km] - 0.5 * slope[v];
                    R[v * n + km] = in[v * n + km] + 0.5 * slope[v];
                }
                L[km] = L[km] - Bl;
                R[km] = R[km] - Br;
            }
    }
    // swe_2D.jl:43-81 and centralupwind.jl:48-77
    void phys(int d, const T& h, const T& hu, const T& hv, T f[3], T e[2]) const {
        const T hs = desing(h);
        const T u = hu / hs, v = hv / hs;
        const T h2 = h * h;
        const T cc = sqrt_(c.g * h);
        const double rho = c.rho, g = c.g;
        if (d == 0) {
            f[0] = rho * h * u; f[1] = rho * h * u * u + 0.5 * rho * g * h2; f[2] = rho * h * u * v;
            e[0] = u + cc; e[1] = u - cc;
        } else {
            f[0] = rho * h * v; f[1] = rho * h * u * v; f[2] = rho * h * v * v + 0.5 * rho * g * h2;
            e[0] = v + cc; e[1] = v - cc;
        }
    }
    T cu(int d, const T m[3], const T q[3], T F[3]) const {
        T fm[3] = {T(0.0), T(0.0), T(0.0)}, em[2] = {T(0.0), T(0.0)}, fp[3] = {T(0.0), T(0.0), T(0.0)},
          ep[2] = {T(0.0), T(0.0)};
        if (val(m[0]) > c.depth_cutoff) phys(d, m[0], m[1], m[2], fm, em);
        if (val(q[0]) > c.depth_cutoff) phys(d, q[0], q[1], q[2], fp, ep);
        const T ap = jmax(jmax(ep[0], em[0]), T(0.0));
        const T am = jmin(jmin(ep[1], em[1]), T(0.0));
        if (std::fabs(val(ap - am)) < c.desingularizing_kappa) { F[0] = F[1] = F[2] = T(0.0); return T(0.0); }
        for (int v = 0; v < 3; ++v) F[v] = (ap * fm[v] - am * fp[v]) / (ap - am) + ((ap * am) / (ap - am)) * (q[v] - m[v]);
        if (val(m[0]) < c.depth_cutoff && val(q[0]) < c.depth_cutoff) return T(0.0);
        return jmax(abs_(ap), abs_(am));
    }
    // numericalflux.jl:21-33 then source_term_bottom.jl:24-41; returns max ws
    T flux_and_bottom(std::vector<T>& o, int d) {
        const double dd = d == 0 ? c.dx : c.dy, inv = 1 / dd;
        const size_t step = d == 0 ? 1 : (size_t)nxT;
        #pragma omp parallel for schedule(static)
        for (int j = c.gc; j < nyT - c.gc; ++j)
            for (int i = c.gc; i < nxT - c.gc; ++i) {
                const size_t km = id(i, j), kl = km - step, kr = km + step;
                T Rm[3], Lr[3], Rl[3], Lm[3], Fr[3], Fl[3];
                for (int v = 0; v < 3; ++v) { Rm[v] = R[v * n + km]; Lr[v] = L[v * n + kr]; Rl[v] = R[v * n + kl]; Lm[v] = L[v * n + km]; }
                const T sr = cu(d, Rm, Lr, Fr), sl = cu(d, Rl, Lm, Fl);
                for (int v = 0; v < 3; ++v) o[v * n + km] = o[v * n + km] - inv * (Fr[v] - Fl[v]);
                ws[km] = jmax(sr, sl);
            }
        T m = ws[0];                                  // maximum(ws): first maximum in memory order
        for (size_t k = 1; k < n; ++k) if (val(ws[k]) > val(m)) m = ws[k];
        if (!c.bottom_source) return m;
        #pragma omp parallel for schedule(static)
        for (int j = c.gc; j < nyT - c.gc; ++j)
            for (int i = c.gc; i < nxT - c.gc; ++i) {
                const size_t k = id(i, j);
                const T hsum = R[k] + L[k];
                o[(1 + d) * n + k] = o[(1 + d) * n + k] + -c.g * ((B_right(i, j, d) - B_left(i, j, d)) / dd) * (hsum / 2.0);
            }
        return m;
    }
    double rain_at(double t) const {
        if (c.rain == SFV_RAIN_CONSTANT) return c.rain_rate / 3600;
        if (c.rain == SFV_RAIN_TIME_TABLE) {
            size_t index = 0;
            while (index + 1 < rain_times.size()) { if (t < rain_times[index + 1]) break; index += 1; }
            return rain_rates[index] / 3600;
        }
        return 0.0;
    }
    void sources(std::vector<T>& o, double t) {
        const double r = rain_at(t);
        const double tf = c.infiltration == SFV_INFILTRATION_HORTON ? c.horton_fc + (c.horton_f0 - c.horton_fc) * std::exp(-c.horton_k * t)
                                                                    : c.infiltration_rate;
        auto rain = [&]() { if (c.rain) for (size_t k = 0; k < n; ++k) o[k] = o[k] + r; };
        auto infil = [&]() {
            if (c.infiltration == SFV_INFILTRATION_HORTON) for (size_t k = 0; k < n; ++k) o[k] = o[k] - factor[k] * tf;
            else if (c.infiltration == SFV_INFILTRATION_CONSTANT) for (size_t k = 0; k < n; ++k) o[k] = o[k] - tf;
        };
        if (c.rain_before_infiltration) { rain(); infil(); } else { infil(); rain(); }
    }
    T friction_law(const T& hs, const T& speed) const {
        if (c.friction == SFV_FRICTION_BSA2012) return -Cz * speed / (cbrt_(hs) * hs);
        if (c.friction == SFV_FRICTION_FCG2016) return -Cz * speed / (cbrt_(hs) * hs * hs);
        return -Cz * speed / (hs * hs);
    }
    void bc(std::vector<T>& Uv) {
        const int gc = c.gc;
        auto put = [&](size_t dst, size_t src, int neg) {
            for (int v = 0; v < 3; ++v) Uv[v * n + dst] = (v == neg) ? -Uv[v * n + src] : Uv[v * n + src];
        };
        for (int j = gc; j < nyT - gc; ++j)
            for (int g = 1; g <= gc; ++g) {
                if (c.bc == SFV_BC_WALL) { put(id(g - 1, j), id(2 * gc - g, j), 1); put(id(nxT - gc + g - 1, j), id(nxT - gc - g, j), 1); }
                else { put(id(g - 1, j), id(nxT + g - 2 * gc - 1, j), -1); put(id(nxT - (gc - g) - 1, j), id(gc + g - 1, j), -1); }
            }
        for (int i = gc; i < nxT - gc; ++i)
            for (int g = 1; g <= gc; ++g) {
                if (c.bc == SFV_BC_WALL) { put(id(i, g - 1), id(i, 2 * gc - g), 2); put(id(i, nyT - gc + g - 1), id(i, nyT - gc - g), 2); }
                else { put(id(i, g - 1), id(i, nyT + g - 2 * gc - 1), -1); put(id(i, nyT - (gc - g) - 1), id(i, gc + g - 1), -1); }
            }
    }
    // simulator.jl:145-177 with max_dt; returns dt
    T step(double max_dt, double t) {
        const int nsub = c.stepper == SFV_RUNGE_KUTTA2 ? 2 : 1;
        for (int s = 0; s < nsub; ++s) {
            std::vector<T>& o = U[s + 1];
            const std::vector<T>& in = U[s];
            for (auto& x : o) x = T(0.0);
            T a[2];
            for (int d = 0; d < 2; ++d) { reconstruct(in, d); a[d] = flux_and_bottom(o, d); }
            sources(o, t);
            if (s == 0) dt_cur = jmin(c.cfl * jmin(c.dx / a[0], c.dy / a[1]), T(max_dt));
            const T dt = dt_cur;
            for (size_t k = 0; k < 3 * n; ++k) o[k] = in[k] + dt * o[k];
            if (nsub == 2 && s == 1) for (size_t k = 0; k < 3 * n; ++k) o[k] = 0.5 * (U[0][k] + o[k]);
            if (c.friction) {
                #pragma omp parallel for schedule(static)
                for (int j = 0; j < nyT; ++j)
                    for (int i = 0; i < nxT; ++i) {
                        const size_t k = id(i, j);
                        const T hs = desing(in[k] - B_cell(i, j));
                        const T u = in[n + k] / hs, v = in[2 * n + k] / hs;
                        T speed = sqrt_(u * u + v * v);
                        if (val(speed) == 0.0) speed = T(0.0);
                        const T fs = friction_law(hs, speed);
                        o[k] = o[k] / (1 - dt * 0.0);
                        o[n + k] = o[n + k] / (1 - dt * fs);
                        o[2 * n + k] = o[2 * n + k] / (1 - dt * fs);
                    }
            }
            for (int j = 0; j < nyT; ++j)
                for (int i = 0; i < nxT; ++i) {
                    const size_t k = id(i, j);
                    const T b = B_cell(i, j);
                    if (val(o[k] - b) < c.depth_cutoff) { o[k] = jmax(o[k], b); o[n + k] = T(0.0); o[2 * n + k] = T(0.0); }
                }
            bc(o);
        }
        std::swap(U[0], U[nsub]);
        return dt_cur;
    }
};

}  // namespace oracle_ad
