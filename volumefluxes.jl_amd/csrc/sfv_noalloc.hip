// sfv_noalloc.hip — sfv_advance_noalloc: the device-resident stepping of
// sfv_advance on buffers the caller owns, with no device allocation.
//
// The kernels of a fused RK2 step (march stage 1, march stage 2 closing the
// step, the boundary condition when the loop ends) need, besides the state
// and the topography the caller passes, three scratch buffers: the stage
// right-hand side (3 planes), the other state of the ping-pong pair and the
// step-control block.  Here they live in one file-scope device array,
// `scratch`, which every GPU the library runs on holds from module load until
// the process ends.  Its device address is looked up once per device with
// hipGetSymbolAddress (the array's name in host code is not a device
// address) and cached; kernels and copies only ever see that address.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/sinfvm_hip.h"
#include "sfv_handle.hpp"

namespace {

constexpr size_t CTL_BYTES = 256;                // StepCtl, padded to the array's alignment
static_assert(sizeof(sfv::StepCtl) <= CTL_BYTES, "StepCtl outgrew its scratch slot");

static __device__ __align__(256) unsigned char scratch[SFV_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};               // device address of `scratch` per device

int fail(int code, const std::string& msg) {
    sfv::set_call_error(msg);          // read with sfv_create_error()
    return code;
}

// Largest nxT*nyT whose right-hand side and second state fit the array
// beside the control block (each rounded up to 256 bytes).
int64_t max_cells(int32_t precision) {
    const size_t elem = precision == SFV_F32_STORAGE ? 4 : 8;
    return (int64_t)((SFV_NOALLOC_SCRATCH_BYTES - CTL_BYTES - 512) / (2 * 3 * elem));
}

size_t round_up(size_t b) { return (b + 255) / 256 * 256; }

// The address of this device's `scratch`, looked up on the device's first call.
int scratch_of(int dev, void** out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SFV_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return SFV_OK;
}

#define NA_TRY(expr)                                                                          \
    do {                                                                                      \
        const hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) return fail(SFV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class S>
int advance(const sfv::KParams& P, hipStream_t st, void* base, const double* B, S* U, int64_t nsteps, double max_dt,
            double t0, double* t_out, double* dt_out) {
    using namespace sfv;
    const size_t n3 = 3 * (size_t)P.nxT * P.nyT;
    StepCtl* ctl = (StepCtl*)base;
    S* RHS = (S*)((unsigned char*)base + CTL_BYTES);
    S* other = (S*)((unsigned char*)RHS + round_up(n3 * sizeof(S)));
    S* buf[2] = {U, other};
    // the step control is armed by a kernel (no memset): every field a step
    // reads is written here, whatever an earlier call left in the array
    NA_TRY(launch_prepare(st, P, ctl, 0, t0, NAN, 0, 1, max_dt, (long long)nsteps, 0,
                          sizeof(S) == 8 ? SFV_F64 : SFV_F32_STORAGE));
    for (int64_t s = 0; s < nsteps; ++s) {
        const int slot = (int)(s & 1);
        S* in = buf[s & 1];
        S* out = buf[(s + 1) & 1];
        NA_TRY(launch_march<S>(st, P, ctl, slot, 1, in, nullptr, RHS, B, 0, 1));
        NA_TRY(launch_march<S>(st, P, ctl, slot, 2, in, RHS, out, B, 1, 0));
    }
    S* last = buf[nsteps & 1];
    NA_TRY(launch_bc<S>(st, P, ctl, (int)(nsteps & 1), last, BC_STANDALONE));   // the march path leaves ghosts
    if (last != U) NA_TRY(launch_copy(st, last, U, n3 * sizeof(S)));
    StepCtl host;
    NA_TRY(hipMemcpyAsync(&host, ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, st));
    NA_TRY(hipStreamSynchronize(st));
    *t_out = host.t;
    if (dt_out) *dt_out = host.dt;
    if (host.steps != nsteps) return fail(SFV_ERR_HIP, "the step loop ended early");
    if (host.nonfinite) return fail(SFV_ERR_NONFINITE_DT, "a time step produced a non-finite dt");
    return SFV_OK;
}

}  // namespace

extern "C" {

int sfv_advance_noalloc(const sfv_config* cfg, const double* B_dev, void* U_dev, int64_t nsteps, double max_dt,
                        double* t_io, double* dt_out, void* stream) {
    sfv::set_call_error("");
    std::string why;
    const int vr = sfv::validate_config(cfg, why);
    if (vr) return fail(vr, why);
    if (!B_dev || !U_dev || !t_io) return fail(SFV_ERR_ARGUMENT, "B_dev, U_dev and t_io must not be NULL");
    if ((uintptr_t)U_dev % 16 || (uintptr_t)B_dev % 8)
        return fail(SFV_ERR_ARGUMENT, "U_dev must be 16-byte aligned and B_dev 8-byte aligned (hipMalloc results are)");
    if (nsteps < 0) return fail(SFV_ERR_ARGUMENT, "nsteps must be >= 0");
    // the fused RK2 kernels: ghost-cell sources inside the grid, sources
    // that need no device table
    if (cfg->stepper != SFV_RUNGE_KUTTA2) return fail(SFV_ERR_ARGUMENT, "sfv_advance_noalloc runs RungeKutta2");
    if (cfg->gc < 2 || cfg->nx < cfg->gc || cfg->ny < cfg->gc || cfg->bc == SFV_BC_NEUMANN)
        return fail(SFV_ERR_ARGUMENT, "sfv_advance_noalloc needs WallBC or PeriodicBC with nx, ny >= gc >= 2");
    if (cfg->rain != SFV_RAIN_NONE && cfg->rain != SFV_RAIN_CONSTANT)
        return fail(SFV_ERR_ARGUMENT, "sfv_advance_noalloc takes no rain table or rain field (ConstantRain only)");
    if (cfg->infiltration == SFV_INFILTRATION_HORTON)
        return fail(SFV_ERR_ARGUMENT, "sfv_advance_noalloc takes no Horton factor field (ConstantInfiltration only)");
    const int64_t cells = (int64_t)(cfg->nx + 2 * cfg->gc) * (cfg->ny + 2 * cfg->gc);
    if (cells > max_cells(cfg->precision))
        return fail(SFV_ERR_DOMAIN, "grid of " + std::to_string(cells) + " cells (ghosts included) exceeds the " +
                                        std::to_string(max_cells(cfg->precision)) +
                                        " cells of the no-allocation scratch (SFV_NOALLOC_SCRATCH_BYTES)");
    if (nsteps == 0) return SFV_OK;
    int dev = 0;
    NA_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(SFV_ERR_HIP, "device index out of range");
    void* base = nullptr;
    const int rs = scratch_of(dev, &base);
    if (rs) return rs;
    sfv::KParams P;
    sfv::build_kparams(*cfg, P, nullptr, 0, nullptr, nullptr, 0);
    hipStream_t st = (hipStream_t)stream;
    if (cfg->precision == SFV_F64)
        return advance<double>(P, st, base, B_dev, (double*)U_dev, nsteps, max_dt, *t_io, t_io, dt_out);
    return advance<float>(P, st, base, B_dev, (float*)U_dev, nsteps, max_dt, *t_io, t_io, dt_out);
}

}  // extern "C"
