// sfv_handle.hpp — the object behind the C ABI's opaque handle: one grid's
// device buffers, its stream and the kernel parameter block.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sinfvm_hip.h"
#include "sfv_pool.hpp"
#include "swe_kernels.hpp"

namespace sfv {
// The reference constructors' checks of a configuration (sfv_capi.hip).
int validate_config(const sfv_config* c, std::string& why);
// The message sfv_create_error() returns on this thread (sfv_capi.hip).
void set_call_error(const std::string& msg);
// bc.jl's ghost-cell assignments replayed on an index map (sfv_capi.hip).
void ghost_maps(int bc, int gc, int N, int* lo, int* hi, int* flo, int* fhi);

// The kernel parameter block of a configuration, pointing at the device
// copies of its tables (rain table: rates then times; Horton factor field;
// FunctionalRain field), each NULL when the configuration has none.
// Allocates nothing.
inline void build_kparams(const sfv_config& c, KParams& p, const double* d_rain, int rain_n, const double* d_factor,
                          const double* d_rain_field, int march_rows) {
    std::memset(&p, 0, sizeof(p));
    p.nx = c.nx; p.ny = c.ny; p.gc = c.gc;
    p.nxT = c.nx + 2 * c.gc; p.nyT = c.ny + 2 * c.gc;
    p.dx = c.dx; p.dy = c.dy;
    p.inv_dx = 1 / c.dx; p.inv_dy = 1 / c.dy;
    auto pow2 = [](double d) { int e; return std::isfinite(d) && d > 0 && std::frexp(d, &e) == 0.5; };
    p.dx_pow2 = pow2(c.dx) ? 1 : 0;
    p.dy_pow2 = pow2(c.dy) ? 1 : 0;
    p.phys.g = c.g; p.phys.rho = c.rho; p.phys.cutoff = c.depth_cutoff; p.phys.kappa = c.desingularizing_kappa;
    p.theta = c.theta; p.cfl = c.cfl; p.Cz = c.Cz;
    p.rec = c.reconstruction == SFV_LINEAR_RECONSTRUCTION ? 1 : 0;
    p.friction = c.friction;
    p.bottom = c.bottom_source ? 1 : 0;
    p.rain_kind = c.rain;
    p.infil_kind = c.infiltration;
    p.rain_first = c.rain_before_infiltration ? 1 : 0;
    p.stepper = c.stepper;
    p.rain_n = rain_n;
    p.rain_rates = d_rain;
    p.rain_times = d_rain ? d_rain + rain_n : nullptr;
    p.rain_const = c.rain_rate / 3600;
    p.infil_rate = c.infiltration_rate;
    p.f0 = c.horton_f0; p.fc = c.horton_fc; p.hk = c.horton_k;
    p.factor = d_factor;
    p.rain_field = d_rain_field;
    p.bcell_1d = c.embedded_1d ? 1 : 0;
    p.bc = c.bc;
    ghost_maps(c.bc, c.gc, p.nxT, p.src_lo_x, p.src_hi_x, p.flip_lo_x, p.flip_hi_x);
    ghost_maps(c.bc, c.gc, p.nyT, p.src_lo_y, p.src_hi_y, p.flip_lo_y, p.flip_hi_y);
    p.rows = march_rows;
    // central_upwind_rng's parameter conditions; SFV_RANGED=0 turns it off
    auto in = [](double v, double lo, double hi) { return v >= lo && v <= hi; };
    const char* rng_env = std::getenv("SFV_RANGED");
    p.ranged = !(rng_env && rng_env[0] == '0') && (p.rec != 1 || in(c.theta, 0x1p-10, 0x1p10)) &&
               in(c.rho, 0x1p-10, 0x1p10) && in(c.g, 0x1p-10, 0x1p10) && in(c.depth_cutoff, 0x1p-39, 0x1p10) &&
               in(c.desingularizing_kappa, 0x1p-39, 0x1p10);
}
}  // namespace sfv

// Kernel variants of an RK2 step (all give identical results).
// Kernel variants of a step.  KV_HYBRID: every right-hand side from the march
// stage-1 kernel, the updates and boundary conditions from the per-pass
// kernels (ForwardEuler and NeumannBC, whose steps the fused stage 2 cannot do).
enum { KV_PER_PASS = 0, KV_TILED = 1, KV_MARCH = 2, KV_HYBRID = 3 };

struct sfv_handle {
    sfv_config cfg{};
    std::vector<double> rain_rates, rain_times, factor;
    sfv::KParams P{};
    int device = 0;
    hipStream_t stream = nullptr;
    size_t elem = 8, ncell = 0;
    // U[0], U[1]: ping-pong states (current = U[cur]); U[2]: the RK2 stage-1
    // state of the per-pass path and a scratch output for kernel timing,
    // allocated on first use.
    void* U[3] = {nullptr, nullptr, nullptr};
    void* RHS = nullptr;               // stage right-hand side, 3 planes
    double* B = nullptr;               // corners (nxT+1, nyT+1)
    double* d_rain = nullptr;          // rates then times
    double* d_factor = nullptr;        // Horton factor field
    double* d_rain_field = nullptr;    // FunctionalRain rates per cell
    sfv::StepCtl* ctl = nullptr;       // device step-control block
    sfv::StepCtl* ctl_host = nullptr;  // pinned mirror
    int64_t cur = 0;                   // index of the current state buffer
    int64_t step_parity = 0;           // parity of the next step's slot
    int64_t device_bytes = 0;
    bool has_state = true;             // the zero-initialised state is a valid state
    bool have_dt = false;
    double last_dt = 0.0;
    int fused = KV_MARCH;               // kernel variant, KV_*
    int march_rows = 0;                 // rows per march segment: 0 = not chosen yet (autotuned on first use)
    bool rows_forced = false;           // set by sfv_set_march_rows
    bool variant_forced = false;        // set by sfv_set_fused: the tuner keeps the variant
    std::string err;
    // Owner-supplied pool (server handles): device blocks (class-rounded),
    // the stream and the pinned mirror come from it and go back to it.
    sfv::DevicePool* pool = nullptr;
    std::vector<std::pair<void*, size_t>> pooled;

    int alloc(void** p, size_t bytes, std::string& why) {
        if (pool) {
            const size_t cls = sfv::DevicePool::size_class(bytes);
            *p = pool->get(cls);
            if (!*p) {
                why = "hipMalloc of " + std::to_string(cls) + " bytes failed";
                return SFV_ERR_OUT_OF_MEMORY;
            }
            pooled.emplace_back(*p, cls);
            device_bytes += (int64_t)cls;
            return SFV_OK;
        }
        if (hipMalloc(p, bytes) != hipSuccess) {
            why = "hipMalloc of " + std::to_string(bytes) + " bytes failed";
            return SFV_ERR_OUT_OF_MEMORY;
        }
        device_bytes += (int64_t)bytes;
        return SFV_OK;
    }

    // Device bytes init() allocates for a configuration (ctl block included;
    // the per-pass scratch state U[2] is not: it is allocated on first use),
    // each buffer rounded to the pool's size class when `pooled_classes`.
    static int64_t init_bytes(const sfv_config& c, bool pooled_classes) {
        const size_t nxT = (size_t)c.nx + 2 * (size_t)c.gc, nyT = (size_t)c.ny + 2 * (size_t)c.gc;
        const size_t es = c.precision == SFV_F32_STORAGE ? 4 : 8, ncell = nxT * nyT;
        std::vector<size_t> sz = {3 * ncell * es, 3 * ncell * es, 3 * ncell * es, (nxT + 1) * (nyT + 1) * 8,
                                  sizeof(sfv::StepCtl)};
        if (c.rain == SFV_RAIN_TIME_TABLE && c.rain_table_len > 0) sz.push_back(2 * (size_t)c.rain_table_len * 8);
        if (c.rain == SFV_RAIN_FIELD) sz.push_back(ncell * 8);
        if (c.infiltration == SFV_INFILTRATION_HORTON) sz.push_back(ncell * 8);
        int64_t b = 0;
        for (size_t x : sz) b += (int64_t)(pooled_classes ? sfv::DevicePool::size_class(x) : x);
        return b;
    }

    int init(const sfv_config& c, std::string& why) {
        cfg = c;
        if (c.rain == SFV_RAIN_TIME_TABLE) {
            rain_rates.assign(c.rain_table_rates, c.rain_table_rates + c.rain_table_len);
            rain_times.assign(c.rain_table_times, c.rain_table_times + c.rain_table_len);
        }
        const int nxT = c.nx + 2 * c.gc, nyT = c.ny + 2 * c.gc;
        ncell = (size_t)nxT * nyT;
        if (c.infiltration == SFV_INFILTRATION_HORTON) {
            if (c.infiltration_factor) factor.assign(c.infiltration_factor, c.infiltration_factor + ncell);
            else factor.assign(ncell, c.infiltration_factor_scalar);
        }
        cfg.rain_table_rates = cfg.rain_table_times = nullptr;
        cfg.infiltration_factor = nullptr;
        elem = c.precision == SFV_F32_STORAGE ? 4 : 8;

        if (hipGetDevice(&device) != hipSuccess) { why = "no HIP device"; return SFV_ERR_HIP; }
        if (pool) stream = pool->get_stream();
        else if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) stream = nullptr;
        if (!stream) { why = "hipStreamCreate failed"; return SFV_ERR_HIP; }
        int rc;
        for (int s = 0; s < 2; ++s)
            if ((rc = alloc(&U[s], 3 * ncell * elem, why))) return rc;
        if ((rc = alloc(&RHS, 3 * ncell * elem, why))) return rc;
        if ((rc = alloc((void**)&B, (size_t)(nxT + 1) * (nyT + 1) * sizeof(double), why))) return rc;
        if ((rc = alloc((void**)&ctl, sizeof(sfv::StepCtl), why))) return rc;
        if (pool) ctl_host = (sfv::StepCtl*)pool->get_pinned(sizeof(sfv::StepCtl));
        else if (hipHostMalloc((void**)&ctl_host, sizeof(sfv::StepCtl), hipHostMallocDefault) != hipSuccess)
            ctl_host = nullptr;
        if (!ctl_host) { why = "hipHostMalloc failed"; return SFV_ERR_OUT_OF_MEMORY; }
        std::memset(ctl_host, 0, sizeof(sfv::StepCtl));
        for (int s = 0; s < 2; ++s) (void)hipMemsetAsync(U[s], 0, 3 * ncell * elem, stream);
        (void)hipMemsetAsync(RHS, 0, 3 * ncell * elem, stream);
        (void)hipMemsetAsync(B, 0, (size_t)(nxT + 1) * (nyT + 1) * sizeof(double), stream);
        (void)hipMemsetAsync(ctl, 0, sizeof(sfv::StepCtl), stream);
        if (!rain_rates.empty()) {
            if ((rc = alloc((void**)&d_rain, 2 * rain_rates.size() * sizeof(double), why))) return rc;
            (void)hipMemcpyAsync(d_rain, rain_rates.data(), rain_rates.size() * sizeof(double), hipMemcpyHostToDevice, stream);
            (void)hipMemcpyAsync(d_rain + rain_rates.size(), rain_times.data(), rain_times.size() * sizeof(double),
                           hipMemcpyHostToDevice, stream);
        }
        if (c.rain == SFV_RAIN_FIELD) {
            if ((rc = alloc((void**)&d_rain_field, ncell * sizeof(double), why))) return rc;
            (void)hipMemsetAsync(d_rain_field, 0, ncell * sizeof(double), stream);
        }
        if (!factor.empty()) {
            if ((rc = alloc((void**)&d_factor, ncell * sizeof(double), why))) return rc;
            (void)hipMemcpyAsync(d_factor, factor.data(), ncell * sizeof(double), hipMemcpyHostToDevice, stream);
        }
        if (hipStreamSynchronize(stream) != hipSuccess) { why = "initialisation failed"; return SFV_ERR_HIP; }
        build_params();
        fused = fusable() ? KV_MARCH : hybrid_ok() ? KV_HYBRID : KV_PER_PASS;
        return SFV_OK;
    }

    // U[2] is allocated on first use: only the per-pass and hybrid RK2 paths (their stage-1
    // state) and kernel timing (a scratch output) need it.
    int ensure_scratch(std::string& why) {
        if (U[2]) return SFV_OK;
        return alloc(&U[2], 3 * ncell * elem, why);
    }

    void build_params() {
        sfv::build_kparams(cfg, P, d_rain, (int)rain_rates.size(), d_factor, d_rain_field, march_rows);
    }

    // The march stage-1 kernel reads a 2-cell halo of the state and writes
    // the interior right-hand side; it needs gc >= 2 and nothing else.
    bool hybrid_ok() const { return P.gc >= 2; }

    // Whether every ghost cell's boundary-condition source is an interior
    // cell (WallBC, PeriodicBC with nx, ny >= gc): kernels may then read a
    // ghost through the maps instead of from memory.
    bool ghost_sources_interior() const {
        if (P.gc < 2) return false;
        for (int g = 0; g < P.gc; ++g) {
            const int sx[2] = {P.src_lo_x[g], P.src_hi_x[g]}, sy[2] = {P.src_lo_y[g], P.src_hi_y[g]};
            for (int k = 0; k < 2; ++k) {
                if (sx[k] < P.gc || sx[k] >= P.nxT - P.gc) return false;
                if (sy[k] < P.gc || sy[k] >= P.nyT - P.gc) return false;
            }
        }
        return true;
    }

    // Does a step leave the output's ghost cells unwritten (they are then
    // filled where the state leaves the stepping loop)?  The march and tiled
    // RK2 paths and the two-launch ForwardEuler path do.
    bool leaves_ghosts() const {
        return fused == KV_MARCH || fused == KV_TILED ||
               (fused == KV_HYBRID && cfg.stepper == SFV_FORWARD_EULER && ghost_sources_interior());
    }

    // The fused stage kernels rebuild stage-1 ghost cells from their
    // boundary-condition sources, which must then be interior cells.
    bool fusable() const { return cfg.stepper == SFV_RUNGE_KUTTA2 && ghost_sources_interior(); }

    ~sfv_handle() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (pool) {
            for (auto& b : pooled) pool->put(b.first, b.second);
            if (ctl_host) pool->put_pinned(ctl_host);
            if (stream) pool->put_stream(stream);
            return;
        }
        for (void* p : U) if (p) (void)hipFree(p);
        if (RHS) (void)hipFree(RHS);
        if (B) (void)hipFree(B);
        if (d_rain) (void)hipFree(d_rain);
        if (d_factor) (void)hipFree(d_factor);
        if (d_rain_field) (void)hipFree(d_rain_field);
        if (ctl) (void)hipFree(ctl);
        if (ctl_host) (void)hipHostFree(ctl_host);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace sfv {
// sfv_create with an optional owner pool (nullptr: plain allocations).
int create_handle(const sfv_config* cfg, DevicePool* pool, sfv_handle** out, std::string& why);
// Choose the march kernels' segment height for a handle (measured on the
// handle's current state, cached per grid size class); a no-op once chosen.
int tune_march_rows(sfv_handle* h, std::string& why);
// The cached height for the handle's size class, without measuring (0: none yet).
int cached_march_rows(sfv_handle* h);
// False when SFV_AUTOTUNE=0 is set in the environment: no height is measured.
bool autotune_enabled();
}  // namespace sfv
