"""The reference's realistic terrain example on the HIP backend: the bay
terrain of examples/terrain.jl:59-136 (bay.txt loaded and coarsened twice
with example_tools.jl:24-45, low entries in-filled as terrain.jl:37-56 does;
the arrays are the committed fixture tests/golden/bay_terrain.npz, written
by tests/golden/make_bay_fixture.py).

1. The example's own configuration (terrain.jl:85-135): 107 x 145 cells over
   the original 451 x 601 extent (dx = 451/107, dy = 601/145: no power-of-two
   spacing), NeumannBC, sources [HortonInfiltration, ConstantRain(15/1000),
   SourceTermBottom] in that order, ImplicitFriction(), depth_cutoff = 8e-2,
   ForwardEuler, a zero initial state, simulate_to_time with
   maximum_timestep = 1.0 and an IntervalWriter(10 s) total-water callback:
   step by step for 60 steps, and the full 120 s with Horton replaced by a
   constant rate (see the tests for why).
2. The same terrain flooded from one side (w = max(B_cell, 6 m) for x < 150 m)
   with RungeKutta2: wet/dry fronts running over real terrain.
3. The example's commented-out alternative ConstantInfiltration(15/1000/3600)
   (terrain.jl:87) with WallBC and RungeKutta2 (the fused march kernels).

Horton's exp is the device libm's against the host's, so the runs with it are
checked to 1e-13 relative per field; the others bitwise.
"""
import os

import numpy as np
import pytest

import oracle as O
import volumefluxes as VF
from volumefluxes.io import TotalWaterVolume

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _terrain():
    d = np.load(os.path.join(HERE, "golden", "bay_terrain.npz"))
    return np.asfortranarray(d["terrain"]), d["upper_corner"]


def _build(bc, stepper, infiltration, U0_fn):
    terrain, upper = _terrain()
    nx, ny = terrain.shape[0] - 5, terrain.shape[1] - 5
    grid = VF.CartesianGrid(nx, ny, gc=2, boundary=bc, extent=[[0.0, float(upper[0])], [0.0, float(upper[1])]])
    bottom = VF.BottomTopography2D(terrain, None, grid)
    eq = VF.ShallowWaterEquations(bottom, depth_cutoff=8e-2)
    infil = VF.HortonInfiltration(grid) if infiltration == "horton" else VF.ConstantInfiltration(float(infiltration))
    system = VF.ConservedSystem(None, VF.LinearReconstruction(), VF.CentralUpwind(eq), eq, grid,
                                [infil, VF.ConstantRain(15 / 1000.0), VF.SourceTermBottom()], VF.ImplicitFriction())
    U0 = U0_fn(grid, terrain)
    cfg = VF.build_config(system, stepper, grid, 0.25, 0)
    o = O.Oracle2D(cfg, terrain)
    o._cfg_keep = cfg
    o.set_interior(U0)
    sim = VF.Simulator(VF.HIPBackend(np.float64), system, stepper, grid)
    VF.set_current_state(sim, U0)
    return grid, bottom, o, sim


def _zero(grid, terrain):
    return np.zeros((grid.nx, grid.ny, 3), order="F")


def _flood(grid, terrain):
    from volumefluxes.scenarios import cell_bottom
    Bc = cell_bottom(terrain)[2:-2, 2:-2]
    x = grid.cell_centers()[:, :, 0]
    U = np.zeros((grid.nx, grid.ny, 3), order="F")
    U[:, :, 0] = np.where(x < 150.0, np.maximum(Bc, 6.0), Bc)
    return U


def _rel(a, b):
    return max(float(np.max(np.abs(a[..., v] - b[..., v])) / max(np.max(np.abs(b[..., v])), 1e-300))
               for v in range(3))


def test_bay_example_configuration_first_steps():
    # The example as written (Horton's exp).  Its wet/dry clamp at the 8 cm
    # cutoff amplifies last-bit differences: the oracle against itself with
    # exp changed by one ulp differs by < 4e-17 over the first 60 steps but by
    # 3e-2 after the full 120 s (10,048 steps).  So the device's exp is checked
    # here step by step, and the full run below without exp.
    grid, bottom, o, sim = _build(VF.NeumannBC(), VF.ForwardEulerStepper(), "horton", _zero)
    assert sim.kernels == "hybrid"
    worst = 0.0
    for k in range(60):
        max_dt = min(1.0, 120.0 - o.t)
        dr = o.step(max_dt, o.t); o.t += dr
        VF.perform_step(sim, max_dt); sim.t += sim.current_timestep
        assert sim.current_timestep == dr, k
        worst = max(worst, _rel(sim.get_full_state(), o.state()))
        assert worst <= 1e-13, k
    print(f"[bay] example config (Horton): 60 steps to t={sim.t:.4f}, max rel {worst:.2e}", flush=True)


def test_bay_example_configuration_full_run_bitwise():
    # The example's configuration with Horton replaced by ConstantInfiltration
    # at Horton's initial rate f0 (no exp): the whole 120 s run, with the
    # example's maximum_timestep and IntervalWriter callback, bitwise.
    grid, bottom, o, sim = _build(VF.NeumannBC(), VF.ForwardEulerStepper(), 1.977e-4, _zero)
    writer = TotalWaterVolume(bottom)
    n = o.simulate_to_time(120.0, maximum_timestep=1.0)
    VF.simulate_to_time(sim, 120.0, maximum_timestep=1.0, callback=[VF.IntervalWriter(step=10.0, writer=writer)])
    S, R = sim.get_full_state(), o.state()
    wet = int(np.sum(S[2:-2, 2:-2, 0] - VF.io.bottom_per_cell(bottom)[2:-2, 2:-2] > 8e-2))
    print(f"[bay] example config (constant f0): {n} steps to t={sim.t}, wet cells {wet}, "
          f"{len(writer.water_volume)} volume samples, max rel {_rel(S, R):.2e}", flush=True)
    assert n >= 50 and wet > 1000
    assert sim.t == o.t == 120.0
    assert len(writer.water_volume) >= 12
    assert np.array_equal(S, R)


def test_bay_flood_runge_kutta2():
    grid, bottom, o, sim = _build(VF.NeumannBC(), VF.RungeKutta2(), "horton", _flood)
    o.advance(80)
    VF.advance_steps(sim, 80)
    S, R = sim.get_full_state(), o.state()
    rel = _rel(S, R)
    moving = int(np.sum(np.abs(S[2:-2, 2:-2, 1]) > 1e-3))
    print(f"[bay] flood RK2: t={sim.t} (oracle {o.t}), moving cells {moving}, max rel {rel:.2e}", flush=True)
    assert moving > 500
    assert abs(sim.t - o.t) <= 1e-13 * o.t
    assert rel <= 1e-13


def test_bay_constant_infiltration_wall_march_bitwise():
    grid, bottom, o, sim = _build(VF.WallBC(), VF.RungeKutta2(), 15 / 1000.0 / 3600.0, _flood)
    o.advance(80)
    VF.advance_steps(sim, 80)
    print(f"[bay] constant infiltration RK2 wall: t={sim.t} kernels={sim.kernels}", flush=True)
    assert sim.t == o.t
    assert np.array_equal(sim.get_full_state(), o.state())
