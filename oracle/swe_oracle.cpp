// swe_oracle.cpp — C API of the CPU ORACLE (test infrastructure).
// Loaded with ctypes by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg only.  See swe_oracle.hpp for what each function restates.
#include "swe_oracle.hpp"
#include "oracle_ad.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace oracle;

namespace {

struct Handle2D {
    int precision;
    Sim2D<double>* d = nullptr;
    Sim2D<float>* f = nullptr;
};

template <class S>
int64_t simulate_to_time(Sim2D<S>& sim, double endtime, int match_endtime, double maximum_timestep,
                         int64_t max_steps, double* t_io) {
    // simulator.jl:179-212 without callbacks; t is stored as realtype
    S t = (S)*t_io;
    int64_t steps = 0;
    while (t < endtime) {
        if (max_steps > 0 && steps >= max_steps) break;
        double max_dt = match_endtime ? endtime - t : INFINITY;
        if (maximum_timestep > 0) max_dt = jl_min(max_dt, maximum_timestep);
        sim.perform_step(max_dt, (double)t);
        t += sim.current_timestep;
        ++steps;
    }
    *t_io = (double)t;
    return steps;
}

}  // namespace

extern "C" {

void oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int oracle_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// ---- scalar functions, for known-answer tests
double oracle_minmod(double a, double b, double c) { return minmod(a, b, c); }
double oracle_jl_max(double a, double b) { return jl_max(a, b); }
double oracle_jl_min(double a, double b) { return jl_min(a, b); }
double oracle_desingularize(double h, double kappa) { return desingularize(h, kappa); }
double oracle_compute_rain(const sfv_config* c, double t) { Params p = make_params(*c); return compute_rain(p, t); }
double oracle_compute_infiltration(const sfv_config* c, double t, int64_t idx) {
    Params p = make_params(*c);
    return compute_infiltration(p, t, (size_t)idx);
}
// One CentralUpwind face evaluation (centralupwind.jl:48-77) in direction d.
double oracle_central_upwind(const sfv_config* c, int d, const double* m, const double* q, double* F) {
    Sim2D<double> sim(*c);
    return sim.central_upwind(d, m, q, F);
}

// Bottom lookups of a 2D handle's topography (bottom_topography.jl:97-101):
// which 0 = B_cell, 1 = B_face_left, 2 = B_face_right; 0-based cell (i, j).
double oracle2d_B_lookup(void* hp, int which, int i, int j, int d);

// ---- 2D simulator
void* oracle2d_create(const sfv_config* cfg) {
    Handle2D* h = new Handle2D;
    h->precision = cfg->precision;
    if (cfg->precision == SFV_F32_STORAGE) h->f = new Sim2D<float>(*cfg);
    else h->d = new Sim2D<double>(*cfg);
    return h;
}
void oracle2d_destroy(void* hp) {
    Handle2D* h = (Handle2D*)hp;
    delete h->d;
    delete h->f;
    delete h;
}
void oracle2d_set_rain_field(void* hp, const double* field) {
    Handle2D* h = (Handle2D*)hp;
    auto& f = h->d ? h->d->p.rain_field : h->f->p.rain_field;
    f.assign(field, field + f.size());
}

void oracle2d_set_topography(void* hp, const double* B) {
    Handle2D* h = (Handle2D*)hp;
    if (h->d) std::memcpy(h->d->B.data(), B, h->d->B.size() * sizeof(double));
    else std::memcpy(h->f->B.data(), B, h->f->B.size() * sizeof(double));
}
// set_current_state!: interior (nx, ny, 3) in Float64, then update_bc!
void oracle2d_set_interior(void* hp, const double* U) {
    Handle2D* h = (Handle2D*)hp;
    if (h->d) h->d->set_interior(U, h->d->p.nx, h->d->p.ny);
    else h->f->set_interior(U, h->f->p.nx, h->f->p.ny);
}
// Full (nxT, nyT, 3) state in realtype, copied verbatim.
void oracle2d_set_full(void* hp, const void* U) {
    Handle2D* h = (Handle2D*)hp;
    if (h->d) std::memcpy(h->d->out[0].data(), U, h->d->out[0].size() * sizeof(double));
    else std::memcpy(h->f->out[0].data(), U, h->f->out[0].size() * sizeof(float));
}
void oracle2d_get_full(void* hp, void* U) {
    Handle2D* h = (Handle2D*)hp;
    if (h->d) std::memcpy(U, h->d->out[0].data(), h->d->out[0].size() * sizeof(double));
    else std::memcpy(U, h->f->out[0].data(), h->f->out[0].size() * sizeof(float));
}
// Substep buffer k (0 = current state, 1 = stage-1 output, ...), and the
// system's left/right face buffers and wavespeeds (which: 0 left, 1 right, 2 ws).
void oracle2d_get_substep(void* hp, int k, void* U) {
    Handle2D* h = (Handle2D*)hp;
    if (h->d) std::memcpy(U, h->d->out[k].data(), h->d->out[k].size() * sizeof(double));
    else std::memcpy(U, h->f->out[k].data(), h->f->out[k].size() * sizeof(float));
}
void oracle2d_get_system_buffer(void* hp, int which, void* U) {
    Handle2D* h = (Handle2D*)hp;
    if (h->d) {
        auto& v = which == 0 ? h->d->left : which == 1 ? h->d->right : h->d->ws;
        std::memcpy(U, v.data(), v.size() * sizeof(double));
    } else {
        auto& v = which == 0 ? h->f->left : which == 1 ? h->f->right : h->f->ws;
        std::memcpy(U, v.data(), v.size() * sizeof(float));
    }
}
// perform_step!(sim, max_dt) with sources at time t; returns current_timestep.
double oracle2d_step(void* hp, double max_dt, double t) {
    Handle2D* h = (Handle2D*)hp;
    return h->d ? h->d->perform_step(max_dt, t) : h->f->perform_step(max_dt, t);
}
double oracle2d_current_timestep(void* hp) {
    Handle2D* h = (Handle2D*)hp;
    return h->d ? (double)h->d->current_timestep : (double)h->f->current_timestep;
}
int64_t oracle2d_simulate_to_time(void* hp, double endtime, int match_endtime, double maximum_timestep,
                                  int64_t max_steps, double* t_io) {
    Handle2D* h = (Handle2D*)hp;
    return h->d ? simulate_to_time(*h->d, endtime, match_endtime, maximum_timestep, max_steps, t_io)
                : simulate_to_time(*h->f, endtime, match_endtime, maximum_timestep, max_steps, t_io);
}
// nsteps of perform_step!(sim, max_dt), t += dt after each.
void oracle2d_advance(void* hp, int64_t nsteps, double max_dt, double* t_io) {
    Handle2D* h = (Handle2D*)hp;
    if (h->d) {
        double t = *t_io;
        for (int64_t s = 0; s < nsteps; ++s) { h->d->perform_step(max_dt, t); t += h->d->current_timestep; }
        *t_io = t;
    } else {
        float t = (float)*t_io;
        for (int64_t s = 0; s < nsteps; ++s) { h->f->perform_step(max_dt, (double)t); t += h->f->current_timestep; }
        *t_io = (double)t;
    }
}

double oracle2d_B_lookup(void* hp, int which, int i, int j, int d) {
    Handle2D* h = (Handle2D*)hp;
    if (h->d) return which == 0 ? h->d->B_cell(i, j) : which == 1 ? h->d->B_face_left(i, j, d) : h->d->B_face_right(i, j, d);
    return which == 0 ? h->f->B_cell(i, j) : which == 1 ? h->f->B_face_left(i, j, d) : h->f->B_face_right(i, j, d);
}

// update_bc! on the current state (bc.jl:29-110).
void oracle2d_apply_bc(void* hp) {
    Handle2D* h = (Handle2D*)hp;
    if (h->d) h->d->update_bc(h->d->out[0]);
    else h->f->update_bc(h->f->out[0]);
}

// ---- 1D simulator (Float64)
void* oracle1d_create(const sfv_config* cfg) { return new Sim1D(*cfg); }
void oracle1d_destroy(void* hp) { delete (Sim1D*)hp; }
void oracle1d_set_topography(void* hp, const double* B) {
    Sim1D* s = (Sim1D*)hp;
    std::memcpy(s->B.data(), B, s->B.size() * sizeof(double));
}
void oracle1d_set_interior(void* hp, const double* U) {
    Sim1D* s = (Sim1D*)hp;
    const int n = s->nxT, nx = s->p.nx, gc = s->p.gc;
    for (int v = 0; v < 2; ++v)
        for (int i = 0; i < nx; ++i) s->out[0][v * n + gc + i] = U[v * nx + i];
    s->update_bc(s->out[0]);
}
void oracle1d_get_full(void* hp, double* U) {
    Sim1D* s = (Sim1D*)hp;
    std::memcpy(U, s->out[0].data(), s->out[0].size() * sizeof(double));
}
double oracle1d_step(void* hp, double max_dt, double t) { return ((Sim1D*)hp)->perform_step(max_dt, t); }
int64_t oracle1d_simulate_to_time(void* hp, double endtime, int match_endtime, double maximum_timestep,
                                  int64_t max_steps, double* t_io) {
    Sim1D* s = (Sim1D*)hp;
    double t = *t_io;
    int64_t steps = 0;
    while (t < endtime) {
        if (max_steps > 0 && steps >= max_steps) break;
        double max_dt = match_endtime ? endtime - t : INFINITY;
        if (maximum_timestep > 0) max_dt = jl_min(max_dt, maximum_timestep);
        s->perform_step(max_dt, t);
        t += s->current_timestep;
        ++steps;
    }
    *t_io = t;
    return steps;
}
double oracle1d_current_timestep(void* hp) { return ((Sim1D*)hp)->current_timestep; }

// ---- forward-mode derivative of a fixed-step run (oracle_ad.hpp)
// J = sum over interior cells of (w - B_cell - target)^2 after nsteps steps
// of perform_step!(sim, Inf) from the interior state U0 (nx, ny, 3); with N
// directions (dB[k] over the corners, dCz[k], dU0[k] over the interior; any
// array may be NULL) returns J and writes the N directional derivatives.
}  // extern "C"
template <int N>
static double ad_run(const sfv_config* cfg, const double* B, const double* U0, int64_t nsteps, const double* target,
                     const double* const* dB, const double* dCz, const double* const* dU0, double* dJ_out,
                     double* U_out, double* dts) {
    using DN = oracle_ad::Dual<N>;
    oracle_ad::Step2D<DN> s(*cfg);
    const size_t nB = (size_t)(s.nxT + 1) * (s.nyT + 1);
    for (size_t k = 0; k < nB; ++k) {
        s.B[k] = DN(B[k]);
        for (int q = 0; q < N; ++q) s.B[k].d[q] = dB[q] ? dB[q][k] : 0.0;
    }
    for (int q = 0; q < N; ++q) s.Cz.d[q] = dCz[q];
    const int nx = cfg->nx, ny = cfg->ny, gc = cfg->gc;
    for (int v = 0; v < 3; ++v)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t src = (size_t)v * nx * ny + (size_t)j * nx + i;
                DN x(U0[src]);
                for (int q = 0; q < N; ++q) x.d[q] = dU0[q] ? dU0[q][src] : 0.0;
                s.U[0][v * s.n + s.id(i + gc, j + gc)] = x;
            }
    s.bc(s.U[0]);
    double t = 0.0;
    for (int64_t k = 0; k < nsteps; ++k) {
        const DN dt = s.step(INFINITY, t);
        if (dts) dts[k] = dt.v;
        t += dt.v;
    }
    DN J(0.0);
    for (int j = gc; j < s.nyT - gc; ++j)
        for (int i = gc; i < s.nxT - gc; ++i) {
            const DN e = (s.U[0][s.id(i, j)] - s.B_cell(i, j)) - target[(size_t)(i - gc) + (size_t)nx * (j - gc)];
            J = J + e * e;
        }
    if (U_out)
        for (size_t k = 0; k < 3 * s.n; ++k) U_out[k] = s.U[0][k].v;
    for (int q = 0; q < N; ++q) dJ_out[q] = J.d[q];
    return J.v;
}

extern "C" {
double oracle_ad_run(const sfv_config* cfg, const double* B, const double* U0, int64_t nsteps, const double* target,
                     const double* dB, double dCz, const double* dU0, double* dJ_out, double* U_out, double* dts) {
    const double* b[1] = {dB};
    const double c[1] = {dCz};
    const double* u[1] = {dU0};
    double out[1];
    const double J = ad_run<1>(cfg, B, U0, nsteps, target, b, c, u, out, U_out, dts);
    if (dJ_out) *dJ_out = out[0];
    return J;
}

// Two directions in one pass.
double oracle_ad_run2(const sfv_config* cfg, const double* B, const double* U0, int64_t nsteps, const double* target,
                      const double* dB1, double dCz1, const double* dU1, const double* dB2, double dCz2,
                      const double* dU2, double* dJ_out) {
    const double* b[2] = {dB1, dB2};
    const double c[2] = {dCz1, dCz2};
    const double* u[2] = {dU1, dU2};
    return ad_run<2>(cfg, B, U0, nsteps, target, b, c, u, dJ_out, nullptr, nullptr);
}

// The same run on plain doubles (pins oracle_ad.hpp against Sim2D).
void oracle_ad_forward(const sfv_config* cfg, const double* B, const double* U0, int64_t nsteps, double* U_out,
                       double* dts) {
    oracle_ad::Step2D<double> s(*cfg);
    const size_t nB = (size_t)(s.nxT + 1) * (s.nyT + 1);
    for (size_t k = 0; k < nB; ++k) s.B[k] = B[k];
    const int nx = cfg->nx, ny = cfg->ny, gc = cfg->gc;
    for (int v = 0; v < 3; ++v)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                s.U[0][v * s.n + s.id(i + gc, j + gc)] = U0[(size_t)v * nx * ny + (size_t)j * nx + i];
    s.bc(s.U[0]);
    double t = 0.0;
    for (int64_t k = 0; k < nsteps; ++k) {
        const double dt = s.step(INFINITY, t);
        if (dts) dts[k] = dt;
        t += dt;
    }
    for (size_t k = 0; k < 3 * s.n; ++k) U_out[k] = s.U[0][k];
}

}  // extern "C"
