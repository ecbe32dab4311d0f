"""The single-grid terrain run at full size, whole run: 500 RK2 steps of the
2048 x 2048 Float64 terrain (seed 2048: 32 bumps, 1e-3 slope, w = max(B, 2 m),
Manning n = 0.03, rain 1e-5 m/s, wall boundaries, kappa 1e-4) on the HIP
backend's device-resident loop (sfv_advance: perform_step!(sim, Inf) with
t += dt, simulator.jl:145-177,205) against the oracle driven the same way.

The gate is the reference's full-run tolerance (1e-9 relative per field); the
assertion is stronger: equal t after every 100 steps and a bitwise state.  The
Float32-storage run (make_cpu_backend(Float32): f64 arithmetic rounded to f32
on store) is checked bitwise for its first 40 steps at the same size.
"""
import time

import numpy as np
import pytest

import oracle as O
import volumefluxes as VF
from volumefluxes import _capi as C
from volumefluxes.scenarios import terrain_workload

pytestmark = pytest.mark.gpu
RK2 = VF.RungeKutta2()


def _pair(n, precision):
    grid, system, realtype, U0, B = terrain_workload(n, precision, 2048)
    cfg = VF.build_config(system, RK2, grid, 0.25, C.SFV_F64 if precision == "f64" else C.SFV_F32_STORAGE)
    o = O.Oracle2D(cfg, B)
    o._cfg_keep = cfg
    o.set_interior(U0)
    sim = VF.Simulator(VF.HIPBackend(realtype), system, RK2, grid, cfl=0.25)
    VF.set_current_state(sim, U0)
    return o, sim


def _rel(a, b):
    """max |a - b| / max |b| per field (the reference's parity measure)."""
    return max(float(np.max(np.abs(a[..., v].astype(np.float64) - b[..., v])) / max(np.max(np.abs(b[..., v])), 1e-300))
               for v in range(3))


@pytest.mark.timeout(900)
def test_terrain_2048_f64_full_run_500_steps():
    o, sim = _pair(2048, "f64")
    t0 = time.perf_counter()
    for chunk in range(5):
        o.advance(100)
        VF.advance_steps(sim, 100)
        S, R = sim.get_full_state(), o.state()
        rel = _rel(S, R)
        print(f"[fullrun] f64 2048^2 steps {100 * (chunk + 1)}: t gpu={sim.t!r} oracle={o.t!r} "
              f"max rel {rel:.3e} ({time.perf_counter() - t0:.0f} s)", flush=True)
        assert rel <= 1e-9, chunk            # the full-run gate
        assert sim.t == o.t, chunk
        assert sim.current_timestep == o.current_timestep, chunk
        assert np.array_equal(S, R), chunk   # and, in fact, bitwise


@pytest.mark.timeout(300)
def test_terrain_2048_f32_storage_first_40_steps():
    o, sim = _pair(2048, "f32")
    for chunk in range(2):
        o.advance(20)
        VF.advance_steps(sim, 20)
        S, R = sim.get_full_state(), o.state()
        print(f"[fullrun] f32 2048^2 steps {20 * (chunk + 1)}: t gpu={sim.t!r} oracle={o.t!r}", flush=True)
        assert S.dtype == np.float32 and R.dtype == np.float32
        assert sim.t == o.t, chunk
        assert np.array_equal(S, R), chunk
