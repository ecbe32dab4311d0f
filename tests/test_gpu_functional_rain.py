"""FunctionalRain on the GPU: the host evaluates the rain function on every
cell centre at each step's start time and the kernels add the rates; the
result must equal the oracle fed the same rates, bit for bit, and the
reference's known answers (test_rain.jl:98-160)."""
import ctypes

import numpy as np
import pytest

import volumefluxes as VF
from volumefluxes import _capi as C

from helpers import make_hip, make_oracle, terrain

pytestmark = pytest.mark.gpu


def _lake(fn, nx=10):
    grid = VF.CartesianGrid(nx, nx, gc=2, boundary=VF.WallBC(), extent=[[0.0, 1000.0], [0.0, 1000.0]])
    eq = VF.ShallowWaterEquations()
    system = VF.ConservedSystem(None, VF.LinearReconstruction(), VF.CentralUpwind(eq), eq, grid,
                                [VF.FunctionalRain(fn, grid)])
    sim = VF.Simulator(VF.make_hip_backend(), system, VF.ForwardEulerStepper(), grid)
    U0 = np.zeros((nx, nx, 3)); U0[:, :, 0] = 1.0
    VF.set_current_state(sim, U0)
    return sim


def test_functional_rain_constant_known_answer():
    sim = _lake(lambda t, x, y: 0.01 / 3600.0)
    VF.simulate_to_time(sim, 3 * 3600.0)
    h = VF.current_interior_state(sim).h
    assert abs(sim.t - 3 * 3600.0) <= 1e-10
    assert np.max(np.abs(h - 1.03)) <= 1e-12


@pytest.mark.parametrize("kernels", ["march", "per-pass"])
def test_functional_rain_terrain_bitwise(kernels):
    sc = terrain(70, 50, seed=41, rain=0.0)
    fn = lambda t, x, y: (1e-4 * (1.0 + np.sin(x / 9.0) * np.cos(y / 7.0)) * (1.0 + t)).astype(float) \
        if isinstance(x, np.ndarray) else 1e-4 * (1.0 + np.sin(x / 9.0) * np.cos(y / 7.0)) * (1.0 + t)
    rain = VF.FunctionalRain(fn, sc.grid)
    sc.system.source_terms.append(rain)
    sim = make_hip(sc, kernels)
    o = make_oracle(sc)
    t = 0.0
    for _ in range(12):
        o.set_rain_field(VF.rain_field_values(rain, sc.grid, t))
        dr = o.step(float("inf"), t)
        dh = VF.perform_step(sim, float("inf"))
        assert dh == dr
        t += dr
        sim.t += dh
    assert np.array_equal(sim.get_full_state(), o.state())


def test_device_loops_refuse_a_rain_field():
    sim = _lake(lambda t, x, y: 0.0)
    t = ctypes.c_double(0.0)
    n = ctypes.c_int64()
    assert C.lib().sfv_run(sim.handle, 10.0, 1, 0.0, 0, ctypes.byref(t), ctypes.byref(n)) == C.SFV_ERR_ARGUMENT
    sc = terrain(40, 40, seed=1)
    sc.system.source_terms = [s for s in sc.system.source_terms if not isinstance(s, VF.ConstantRain)]
    sc.system.source_terms.append(VF.FunctionalRain(lambda t, x, y: 0.0, sc.grid))
    with pytest.raises(VF.ArgumentError):
        VF.Ensemble([make_hip(sc)])
