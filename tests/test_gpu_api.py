"""HIP backend through the reference-style API: device-resident
simulate_to_time against the oracle, checkpoint/restore, boundary maps,
sources, NaN and all-dry semantics, and the 2048^2 workload at full size
(bitwise for the first steps; mass conservation as a size-independent
property)."""
import ctypes
import warnings

import numpy as np
import pytest

import volumefluxes as VF
from volumefluxes import _capi as C

from helpers import Scenario, dam_break, make_hip, make_oracle, terrain
from oracle_runs import oracle_2d, system_2d

pytestmark = pytest.mark.gpu
RK2, FE = VF.RungeKutta2(), VF.ForwardEulerStepper()


def _same(a, b):
    """Bitwise equality, with NaN positions compared as positions."""
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na], b[~nb])


def _sim(grid, system, stepper, U0, cfl=0.25, realtype=np.float64, t0=0.0):
    sim = VF.Simulator(VF.HIPBackend(realtype), system, stepper, grid, cfl=cfl, t0=t0)
    VF.set_current_state(sim, U0)
    return sim


@pytest.mark.parametrize("match", [True, False])
def test_simulate_to_time_device_loop_matches_oracle(match):
    sc = terrain(90, 70, seed=31)
    o = make_oracle(sc)
    sim = make_hip(sc)
    o.simulate_to_time(3.0, match_endtime=match)
    VF.simulate_to_time(sim, 3.0, match_endtime=match)
    print(f"[api] t={sim.t} dt={sim.current_timestep}", flush=True)
    assert sim.t == o.t and sim.current_timestep == o.current_timestep
    assert np.array_equal(sim.get_full_state(), o.state())


def test_maximum_timestep_and_callback_paths_agree():
    sc = terrain(50, 40, seed=33)
    o = make_oracle(sc)
    nsteps = o.simulate_to_time(2.0, maximum_timestep=0.01)
    a = make_hip(sc)
    VF.simulate_to_time(a, 2.0, maximum_timestep=0.01)
    times = []
    b = make_hip(sc)
    VF.simulate_to_time(b, 2.0, maximum_timestep=0.01, callback=lambda t, s: times.append(t))
    assert a.t == b.t == o.t == 2.0
    assert np.array_equal(a.get_full_state(), o.state())
    assert np.array_equal(b.get_full_state(), o.state())
    assert len(times) == nsteps and times[-1] == 2.0


def test_interval_writer_callback():
    sc = terrain(40, 30, seed=35)
    sim = make_hip(sc)
    seen = []
    VF.simulate_to_time(sim, 1.0, callback=VF.IntervalWriter(0.25, lambda t, s: seen.append(t)))
    assert len(seen) >= 4


def test_checkpoint_restore_continues_identically():
    sc = terrain(64, 48, seed=37)
    a = make_hip(sc)
    VF.advance_steps(a, 10)
    snap, t_snap = a.get_full_state(), a.t
    VF.advance_steps(a, 15)
    b = make_hip(sc)
    b.set_full_state(snap)
    b.t = t_snap
    VF.advance_steps(b, 15)
    assert a.t == b.t and np.array_equal(a.get_full_state(), b.get_full_state())


def test_set_state_applies_boundary_like_the_oracle():
    for bc in (VF.WallBC(), VF.PeriodicBC(), VF.NeumannBC()):
        sc = terrain(20, 14, seed=39, bc=bc)
        o = make_oracle(sc)
        sim = make_hip(sc)
        assert np.array_equal(sim.get_full_state(), o.state()), type(bc).__name__


def test_interior_layout_roundtrip():
    sc = terrain(33, 21, seed=41)
    sim = make_hip(sc)
    st = VF.current_interior_state(sim)
    assert np.array_equal(st.h, sc.U0[:, :, 0]) and np.array_equal(st.hu, sc.U0[:, :, 1])


def test_rain_and_time_table_bitwise():
    grid = VF.CartesianGrid(10, 10, gc=2, boundary=VF.WallBC(), extent=[[0.0, 1000.0], [0.0, 1000.0]])
    U0 = np.zeros((10, 10, 3)); U0[:, :, 0] = 1.0
    rain = VF.TimeDependentRain(np.array([15.0, 0.3, 20.0]) * 0.001, np.array([0.0, 1.0, 2.0]) * 3600)
    system = system_2d(grid, sources=[rain])
    o = oracle_2d(grid, system, FE, U0)
    sim = _sim(grid, system, FE, U0)
    o.simulate_to_time(3 * 3600)
    VF.simulate_to_time(sim, 3 * 3600)
    assert sim.t == o.t and np.array_equal(sim.get_full_state(), o.state())
    w = VF.current_interior_state(sim).h
    assert np.max(np.abs(w - 1.0353)) <= 1e-4                     # test_rain.jl:94-96


def test_horton_infiltration_within_exp_ulps():
    # exp() is the device libm's, not the host's: ulp-level agreement
    grid = VF.CartesianGrid(10, 10, gc=2, extent=[[0.0, 100.0], [0.0, 100.0]])
    inf = VF.HortonInfiltration(grid, factor=np.linspace(0.5, 1.5, 196).reshape(14, 14))
    U0 = np.zeros((10, 10, 3)); U0[:, :, 0] = 1.0
    system = system_2d(grid, sources=[VF.ConstantRain(0.01), inf])
    o = oracle_2d(grid, system, FE, U0)
    sim = _sim(grid, system, FE, U0)
    o.simulate_to_time(500.0)
    VF.simulate_to_time(sim, 500.0)
    a, b = sim.get_full_state(), o.state()
    assert np.max(np.abs(a - b)) / np.max(np.abs(b)) <= 1e-13


@pytest.mark.parametrize("law", [VF.friction_bsa2012, VF.friction_fcg2016])
def test_other_friction_laws_within_cbrt_ulps(law):
    grid = VF.CartesianGrid(48, 40, gc=2, boundary=VF.WallBC(), extent=[[0.0, 48.0], [0.0, 40.0]])
    sc = terrain(48, 40, seed=43)
    system = system_2d(grid, sc.B, sources=[VF.SourceTermBottom()], kappa=1e-4,
                       friction=VF.ImplicitFriction(Cz=0.02 ** 2, friction_function=law))
    o = oracle_2d(grid, system, RK2, sc.U0)
    sim = _sim(grid, system, RK2, sc.U0)
    for _ in range(10):
        dr = o.step(float("inf"), o.t); o.t += dr
        VF.perform_step(sim, float("inf")); sim.t += sim.current_timestep
    a, b = sim.get_full_state(), o.state()
    assert np.max(np.abs(a - b)) / np.max(np.abs(b)) <= 1e-12


@pytest.mark.parametrize("wet", [True, False])
def test_nan_propagates_like_the_reference(wet):
    # A NaN momentum in a wet cell reaches the wave speeds, so dt and then the
    # whole state become NaN; in a dry cell it only spreads through fluxes.
    sc = terrain(30, 20, seed=45)
    from volumefluxes.scenarios import cell_bottom
    depth = sc.U0[:, :, 0] - cell_bottom(sc.B)[2:-2, 2:-2]
    i, j = np.argwhere(depth > 0.5)[0] if wet else np.argwhere(depth == 0.0)[0]
    U0 = sc.U0.copy()
    U0[i, j, 1] = np.nan
    sc2 = Scenario(sc.name, sc.grid, sc.system, sc.stepper, sc.cfl, U0, sc.realtype, sc.B)
    o = make_oracle(sc2)
    sim = make_hip(sc2)
    for _ in range(3):
        dr = o.step(float("inf"), o.t); o.t += dr
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            dh = VF.perform_step(sim, float("inf"))
        sim.t += dh
        assert (np.isnan(dr) and np.isnan(dh)) or dr == dh
        assert _same(sim.get_full_state(), o.state())
    assert np.isnan(sim.get_full_state()).any()


def test_all_dry_domain_jumps_to_endtime():
    # simulator.jl:190-192: an all-dry domain gives dt = Inf, and with
    # match_endtime the single step lands on t_end
    grid = VF.CartesianGrid(12, 10, gc=2, boundary=VF.WallBC())
    system = system_2d(grid)
    U0 = np.zeros((12, 10, 3))
    o = oracle_2d(grid, system, RK2, U0)
    sim = _sim(grid, system, RK2, U0)
    o.simulate_to_time(5.0)
    VF.simulate_to_time(sim, 5.0)
    assert sim.t == o.t == 5.0
    assert _same(sim.get_full_state(), o.state())


def test_topography_errors_from_the_c_abi():
    sc = terrain(8, 8, seed=47)
    sim = make_hip(sc)
    B = np.zeros((5, 5))
    with pytest.raises(VF.DomainError):
        C.check(sim.handle, C.lib().sfv_set_topography(sim.handle, C.dptr(B), 5, 5))
    # non-zero B without SourceTermBottom: ArgumentError at the boundary too
    grid = VF.CartesianGrid(8, 8, gc=2)
    eq = VF.ShallowWaterEquations()
    system = VF.ConservedSystem(None, VF.LinearReconstruction(), VF.CentralUpwind(eq), eq, grid, [])
    plain = VF.Simulator(VF.HIPBackend(), system, RK2, grid)
    Bn = np.ones((13, 13), order="F")
    with pytest.raises(VF.ArgumentError):
        C.check(plain.handle, C.lib().sfv_set_topography(plain.handle, C.dptr(Bn), 13, 13))


def test_lake_at_rest_on_the_gpu():
    grid = VF.CartesianGrid(124, 124, gc=2, boundary=VF.WallBC(),
                            extent=[[-2 * np.pi, 2 * np.pi], [-2 * np.pi, 2 * np.pi]])
    f = grid.cell_faces(interior=False)
    B = np.cos(f[:, :, 0] + f[:, :, 1]) - 0.5 - 1.5 * (np.hypot(f[:, :, 0], f[:, :, 1]) < 1.0)
    system = system_2d(grid, B, sources=[VF.SourceTermBottom()], rec=VF.LinearReconstruction(2))
    U0 = np.zeros((124, 124, 3)); U0[:, :, 0] = 0.7
    for stepper in (FE, RK2):
        sim = _sim(grid, system, stepper, U0, cfl=0.2)
        VF.simulate_to_time(sim, 0.01)
        st = VF.current_interior_state(sim)
        assert np.max(np.abs(st.hu)) <= 1e-14 and np.max(np.abs(st.hv)) <= 1e-14
        assert np.max(np.abs(st.h - 0.7)) <= 1e-14


# ------------------------------------------------------- full-size workload
def test_terrain_2048_first_steps_bitwise():
    sc = terrain(2048, 2048, seed=2048, width=(20.0, 200.0))
    print("[api] 2048^2 terrain: 10 steps on both", flush=True)
    o = make_oracle(sc)
    sim = make_hip(sc)
    for k in range(10):
        dr = o.step(float("inf"), o.t); o.t += dr
        dh = VF.perform_step(sim, float("inf")); sim.t += dh
        assert dh == dr, k
    assert np.array_equal(sim.get_full_state(), o.state())


def test_mass_conservation_2048_wall_basin():
    n = 2048
    grid = VF.CartesianGrid(n, n, gc=2, boundary=VF.WallBC(), extent=[[0.0, float(n)], [0.0, float(n)]])
    f = grid.cell_faces(interior=False)
    B = 0.5 * np.sin(f[:, :, 0] / 90.0) * np.cos(f[:, :, 1] / 70.0)
    c = grid.cell_centers()
    U0 = np.zeros((n, n, 3))
    U0[:, :, 0] = 2.0 + np.exp(-((c[:, :, 0] - 900) ** 2 + (c[:, :, 1] - 1100) ** 2) / 5e4)
    system = system_2d(grid, B, sources=[VF.SourceTermBottom()])
    sim = _sim(grid, system, RK2, U0)
    m0 = np.sum(VF.current_interior_state(sim).h)
    VF.advance_steps(sim, 100)
    m1 = np.sum(VF.current_interior_state(sim).h)
    print(f"[api] mass drift {abs(m1 - m0) / m0:.3e}", flush=True)
    assert abs(m1 - m0) / m0 <= 1e-12


def test_float32_storage_512_bitwise():
    sc = terrain(512, 512, seed=51, realtype=np.float32, width=(10.0, 60.0))
    o = make_oracle(sc)
    sim = make_hip(sc)
    o.advance(20)
    VF.advance_steps(sim, 20)
    assert sim.t == o.t and np.array_equal(sim.get_full_state(), o.state())
