"""Integration pins of the CPU oracle against the reference's own tests and
against closed forms (file:line cited per test).  Together with the
known-answer tests these are what tie the oracle to the reference; the HIP
backend is then tied to the oracle bitwise (tests/test_gpu_*.py)."""
import math

import numpy as np
import pytest

import oracle as O
import volumefluxes as VF
from volumefluxes import _capi as C

from oracle_runs import (centres_1d, config_1d, dam_break_exact, faces_1d, oracle_2d, system_2d)

FE, RK2 = VF.ForwardEulerStepper(), VF.RungeKutta2()


# ---------------------------------------------------------------- lake at rest
# test/test_lake_at_rest_2d.jl:28-120: step and bumpy bottoms, w0 = 0.7,
# LinearReconstruction(2), ForwardEuler, CFL 0.2, t = 0.01, WallBC.
def _lake_2d(grid, B):
    system = system_2d(grid, B, sources=[VF.SourceTermBottom()], rec=VF.LinearReconstruction(2))
    U0 = np.zeros((grid.nx, grid.ny, 3))
    U0[:, :, 0] = 0.7
    o = oracle_2d(grid, system, FE, U0, cfl=0.2)
    o.simulate_to_time(0.01)
    U = o.state()[2:-2, 2:-2]
    return U


def test_lake_at_rest_2d_step():
    grid = VF.CartesianGrid(64, 64, gc=2, boundary=VF.WallBC(), extent=[[0.0, 10.0], [0.0, 10.0]])
    f = grid.cell_faces(interior=False)
    B = np.where((f[:, :, 0] < 5.0) & (f[:, :, 1] < 5.0), 0.45, 0.55)
    U = _lake_2d(grid, B)
    assert np.max(np.abs(U[:, :, 1])) <= 1e-14
    assert np.max(np.abs(U[:, :, 2])) <= 1e-14
    assert np.max(np.abs(U[:, :, 0] - 0.7)) <= 1e-14


def test_lake_at_rest_2d_bumpy():
    grid = VF.CartesianGrid(124, 124, gc=2, boundary=VF.WallBC(),
                            extent=[[-2 * np.pi, 2 * np.pi], [-2 * np.pi, 2 * np.pi]])
    f = grid.cell_faces(interior=False)
    B = np.cos(f[:, :, 0] + f[:, :, 1]) - 0.5 - 1.5 * (np.hypot(f[:, :, 0], f[:, :, 1]) < 1.0)
    U = _lake_2d(grid, B)
    assert np.max(np.abs(U[:, :, 1])) <= 1e-14
    assert np.max(np.abs(U[:, :, 0] - 0.7)) <= 1e-14


# test/test_lake_at_rest.jl:90-106 (1D): step bottom with WallBC, bumpy bottom
# with PeriodicBC on 1024 cells.
@pytest.mark.parametrize("case", ["step", "bumpy"])
def test_lake_at_rest_1d(case):
    if case == "step":
        nx, x0, x1, bc = 64, 0.0, 10.0, C.SFV_BC_WALL
        f = faces_1d(nx, x0, x1, 2, interior=False)
        B = np.where(f < 5.0, 0.45, 0.55)
    else:
        nx, x0, x1, bc = 1024, -2 * np.pi, 2 * np.pi, C.SFV_BC_PERIODIC
        f = faces_1d(nx, x0, x1, 2, interior=False)
        B = np.cos(f) - 0.5 - 1.5 * (np.abs(f) < 1.0)
    cfg = config_1d(nx, x1, x0=x0, bc=bc, theta=2.0, stepper=C.SFV_FORWARD_EULER, cfl=0.2, bottom=1)
    o = O.Oracle1D(cfg, B)
    U0 = np.zeros((nx, 2))
    U0[:, 0] = 0.7
    o.set_interior(U0)
    o.simulate_to_time(0.01)
    U = o.state()[2:-2]
    assert np.max(np.abs(U[:, 1])) <= 1e-14
    assert np.max(np.abs(U[:, 0] - 0.7)) <= 1e-14


# ----------------------------------------------------------- mirror symmetry
# test/test_shallow_water_dry_states_2d.jl:31-116: a wet disc w = 1 - sign(r^2 - 0.01)
# on a dry flat bottom, 64x64, RK2, periodic, T = 0.05; symmetric to 1e-12.
def test_dry_disc_symmetry():
    nx = ny = 64
    grid = VF.CartesianGrid(nx, ny, gc=2)
    c = grid.cell_centers()
    r2 = (c[:, :, 0] - 0.5) ** 2 + (c[:, :, 1] - 0.5) ** 2
    U0 = np.zeros((nx, ny, 3))
    U0[:, :, 0] = 1 - np.sign(r2 - 0.01)
    o = oracle_2d(grid, system_2d(grid), RK2, U0)
    o.simulate_to_time(0.05)
    assert o.t == 0.05
    U = o.state()[2:-2, 2:-2]
    h, hu, hv = U[:, :, 0], U[:, :, 1], U[:, :, 2]
    assert np.isfinite(U).all()
    tol = 1e-12
    xl, xr = nx // 3 - 1, nx - nx // 3
    yl, yr = ny // 3 - 1, ny - ny // 3
    assert np.max(np.abs(h[xl, :] - h[xr, :])) <= tol
    assert np.max(np.abs(h[:, yl] - h[:, yr])) <= tol
    assert np.max(np.abs(hu[xl, :] + hu[xr, :])) <= tol
    assert np.max(np.abs(hu[:, yl] - hu[:, yr])) <= tol
    assert np.max(np.abs(hv[xl, :] - hv[xr, :])) <= tol
    assert np.max(np.abs(hv[:, yl] + hv[:, yr])) <= tol
    # and the full mirror images
    assert np.max(np.abs(h - h[::-1, :])) <= tol and np.max(np.abs(h - h[:, ::-1])) <= tol
    assert np.max(np.abs(h - h.T)) <= tol


# ------------------------------------------ dimension consistency (bitwise)
# test/test_swashes.jl:218-328: the dam breaks SWASHES 4.1.1 (wet, t = 6) and
# 4.1.2 (dry, t = 4) on 512 cells, kappa = cutoff = 1e-7, CFL 0.1: 2D on
# (n, 3) == 2D on (3, n) transposed == 1D, bitwise, with equal dt.
@pytest.mark.parametrize("hr,t_end", [(0.001, 6.0), (0.0, 4.0)])
@pytest.mark.parametrize("stepper", ["FE", "RK2"])
def test_swashes_1d_2d_bitwise(hr, t_end, stepper):
    n, L, hl = 512, 10.0, 0.005
    st = FE if stepper == "FE" else RK2
    gx = VF.CartesianGrid(n, 3, gc=2, boundary=VF.WallBC(), extent=[[0.0, L], [0.0, 5.0]])
    gy = VF.CartesianGrid(3, n, gc=2, boundary=VF.WallBC(), extent=[[0.0, 5.0], [0.0, L]])
    xs = gx.cell_centers()[:, 0, 0]
    w = np.where(xs < 5.0, hl, hr)
    Ux = np.zeros((n, 3, 3)); Ux[:, :, 0] = w[:, None]
    Uy = np.zeros((3, n, 3)); Uy[:, :, 0] = w[None, :]
    kw = dict(cutoff=1e-7, kappa=1e-7)
    ox = oracle_2d(gx, system_2d(gx, **kw), st, Ux, cfl=0.1)
    oy = oracle_2d(gy, system_2d(gy, **kw), st, Uy, cfl=0.1)
    cfg = config_1d(n, L, stepper=C.SFV_FORWARD_EULER if stepper == "FE" else C.SFV_RUNGE_KUTTA2, cfl=0.1,
                    cutoff=1e-7, kappa=1e-7)
    o1 = O.Oracle1D(cfg, np.zeros(n + 5))
    o1.set_interior(np.stack([np.where(centres_1d(n, 0.0, L) < 5.0, hl, hr), np.zeros(n)], axis=-1))
    for o in (ox, oy, o1):
        o.simulate_to_time(t_end)
    X, Y, D = ox.state()[2:-2, 2:-2], oy.state()[2:-2, 2:-2], o1.state()[2:-2]
    assert np.all(X[:, :, 2] == 0) and np.all(Y[:, :, 1] == 0)
    assert np.array_equal(Y[1, :, 0], X[:, 1, 0]) and np.array_equal(Y[1, :, 2], X[:, 1, 1])
    assert np.array_equal(D[:, 0], X[:, 1, 0]) and np.array_equal(D[:, 1], X[:, 1, 1])
    assert ox.current_timestep == o1.current_timestep == oy.current_timestep


# -------------------------------------------------------- analytic solutions
def _dam_break_1d(n, L, x0, hl, hr, t, stepper=C.SFV_RUNGE_KUTTA2, cfl=0.25, cutoff=1e-5, kappa=1e-5):
    cfg = config_1d(n, L, stepper=stepper, cfl=cfl, cutoff=cutoff, kappa=kappa)
    o = O.Oracle1D(cfg, np.zeros(n + 5))
    xc = centres_1d(n, 0.0, L)
    o.set_interior(np.stack([np.where(xc < x0, hl, hr), np.zeros(n)], axis=-1))
    o.simulate_to_time(t)
    h = o.state()[2:-2, 0]
    he, _ = dam_break_exact(xc, t, x0, hl, hr)
    return np.sum(np.abs(h - he)) * (L / n), o


def test_stoker_dam_break_1024():
    """The 1D CPU workload: 1024 cells on [0, 1000], h = 2 | 1, RK2, CFL 0.25,
    t = 20 s, against the Stoker solution (rarefaction head ~411 m, shock
    ~585 m), converging under refinement."""
    e1024, o = _dam_break_1d(1024, 1000.0, 500.0, 2.0, 1.0, 20.0)
    e512, _ = _dam_break_1d(512, 1000.0, 500.0, 2.0, 1.0, 20.0)
    e256, _ = _dam_break_1d(256, 1000.0, 500.0, 2.0, 1.0, 20.0)
    assert o.t == 20.0
    assert e1024 / 1000.0 < 2e-3            # mean |h - h_exact| below 2 mm on a 1 m jump
    assert e256 > e512 > e1024 and e256 / e1024 > 2.0
    # mass: walls untouched at t = 20 s, so the total water is the initial one
    h = o.state()[2:-2, 0]
    assert abs(h.sum() * (1000.0 / 1024) - (500 * 2.0 + 500 * 1.0)) < 1e-9


@pytest.mark.parametrize("hr,t", [(0.001, 6.0), (0.0, 4.0)])
def test_swashes_dam_breaks_converge(hr, t):
    """SWASHES 4.1.1 (Stoker) and 4.1.2 (Ritter, dry) on [0, 10], x0 = 5,
    hl = 0.005: L1 error of h shrinks under refinement (test/swashes.jl:130-185)."""
    errs = [_dam_break_1d(n, 10.0, 5.0, 0.005, hr, t, stepper=C.SFV_FORWARD_EULER, cfl=0.1, cutoff=1e-7,
                          kappa=1e-7)[0] for n in (128, 256, 512)]
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] / 10.0 < 5e-5            # mean error below 1% of hl


# ------------------------------------------------------- rain / infiltration
# test/test_rain.jl:29-73: 10x10 flat lake at w = 1, WallBC, ForwardEuler.
def _rain_run(rain, hours):
    grid = VF.CartesianGrid(10, 10, gc=2, boundary=VF.WallBC(), extent=[[0.0, 1000.0], [0.0, 1000.0]])
    U0 = np.zeros((10, 10, 3)); U0[:, :, 0] = 1.0
    o = oracle_2d(grid, system_2d(grid, sources=[rain]), FE, U0)
    o.simulate_to_time(hours * 3600)
    assert abs(o.t - hours * 3600) <= 1e-10
    return o.state()[2:-2, 2:-2]


def test_constant_rain_mass():
    U = _rain_run(VF.ConstantRain(0.01), 3)
    assert np.max(np.abs(U[:, :, 0] - (3 * 0.01 + 1.0))) <= 1e-12
    assert np.all(U[:, :, 1] == 0.0) and np.all(U[:, :, 2] == 0.0)


@pytest.mark.parametrize("hours,want,tol", [(1, 0.015 + 1.0, 1e-12), (2, 0.0153 + 1.0, 1e-4),
                                            (3, 0.0353 + 1.0, 1e-4)])
def test_time_dependent_rain(hours, want, tol):
    rain = VF.TimeDependentRain(np.array([15.0, 0.3, 20.0]) * 0.001, np.array([0.0, 1.0, 2.0]) * 3600)
    U = _rain_run(rain, hours)
    assert np.max(np.abs(U[:, :, 0] - want)) <= tol


# test/test_infiltration.jl:26-72: Horton defaults on a 10x10 periodic lake,
# t0 = 1e6 (the rate has decayed to fc), 1000 s: w = 1 - fc*1000.
def test_horton_infiltration_closed_form():
    grid = VF.CartesianGrid(10, 10, gc=2, extent=[[0.0, 100.0], [0.0, 100.0]])
    inf = VF.HortonInfiltration(grid)
    U0 = np.zeros((10, 10, 3)); U0[:, :, 0] = 1.0
    o = oracle_2d(grid, system_2d(grid, sources=[inf]), FE, U0, t0=1e6)
    o.simulate_to_time(1e6 + 1000)
    U = o.state()[2:-2, 2:-2]
    assert np.max(np.abs(U[:, :, 0] - (1.0 - inf.fc * 1000))) <= 1e-10
    assert np.all(U[:, :, 1] == 0.0) and np.all(U[:, :, 2] == 0.0)


# ------------------------------------------------------------ conservation
def _wet_basin(nx=48, ny=40):
    grid = VF.CartesianGrid(nx, ny, gc=2, boundary=VF.WallBC(), extent=[[0.0, 48.0], [0.0, 40.0]])
    f = grid.cell_faces(interior=False)
    B = 0.2 * np.sin(f[:, :, 0] / 7.0) * np.cos(f[:, :, 1] / 5.0)
    c = grid.cell_centers()
    U0 = np.zeros((nx, ny, 3))
    U0[:, :, 0] = 1.0 + 0.3 * np.exp(-((c[:, :, 0] - 20) ** 2 + (c[:, :, 1] - 22) ** 2) / 30.0)
    return grid, B, U0


def test_mass_conservation_wall_basin():
    """With walls and no sources the mass flux through a wall face is exactly
    0 (a+ = -a-), so sum(w) dx dy is conserved to round-off."""
    grid, B, U0 = _wet_basin()
    o = oracle_2d(grid, system_2d(grid, B, sources=[VF.SourceTermBottom()]), RK2, U0)
    m0 = o.state()[2:-2, 2:-2, 0].sum()
    o.advance(200)
    m1 = o.state()[2:-2, 2:-2, 0].sum()
    assert abs(m1 - m0) / m0 <= 1e-13


def test_rain_adds_rate_times_dt_per_step():
    grid, B, U0 = _wet_basin()
    rain = VF.ConstantRain(36.0)        # 1e-2 m/s
    o = oracle_2d(grid, system_2d(grid, B, sources=[VF.SourceTermBottom(), rain]), RK2, U0)
    for _ in range(20):
        m0 = o.state()[2:-2, 2:-2, 0].sum()
        dt = o.step(float("inf"), o.t)
        o.t += dt
        m1 = o.state()[2:-2, 2:-2, 0].sum()
        # RK2 adds rate*dt in each stage and averages: rate*dt per step
        assert abs((m1 - m0) - grid.nx * grid.ny * (36.0 / 3600) * dt) <= 1e-11 * m0


# --------------------------------------------------------------- steppers
def test_rk2_is_heun_average_of_two_euler_steps():
    """RK2 (rungekutta.jl:26-52) is U^{n+1} = (U^n + E(E(U^n)))/2 with E the
    Euler step, for a fixed dt and no friction / clamp activity."""
    grid, B, U0 = _wet_basin()
    sysm = system_2d(grid, B, sources=[VF.SourceTermBottom()])
    dt = 0.01
    rk = oracle_2d(grid, sysm, RK2, U0)
    fe = oracle_2d(grid, sysm, FE, U0)
    Ustart = rk.state()
    assert rk.step(dt, 0.0) == dt
    fe.step(dt, 0.0)
    assert fe.current_timestep == dt
    fe.step(dt, 0.0)
    heun = 0.5 * (Ustart + fe.state())
    assert np.array_equal(rk.state()[2:-2, 2:-2], heun[2:-2, 2:-2])


@pytest.mark.parametrize("stepper,order", [(FE, 1.0), (RK2, 2.0)])
def test_temporal_convergence_order(stepper, order):
    """test/test_timesteppers.jl:76-100 fits the error order of FE and RK2; here
    on a smooth, wet, periodic gravity wave with the spatial scheme fixed
    (NoReconstruction, so the right-hand side is smooth) and dt halved."""
    nx, ny = 64, 2
    grid = VF.CartesianGrid(nx, ny, gc=2, extent=[[0.0, 1.0], [0.0, 1.0]])
    c = grid.cell_centers()
    U0 = np.zeros((nx, ny, 3))
    U0[:, :, 0] = 1.0 + 0.05 * np.sin(2 * np.pi * c[:, :, 0])
    sysm = system_2d(grid, rec=VF.NoReconstruction())
    T = 0.05

    def run(dt):
        o = oracle_2d(grid, sysm, stepper, U0, cfl=10.0)
        o.simulate_to_time(T, maximum_timestep=dt)
        return o.state()[2:-2, 2:-2]

    ref = run(T / 4096)
    dts = [T / 32, T / 64, T / 128, T / 256]
    errs = [np.max(np.abs(run(dt) - ref)) for dt in dts]
    slope = np.polyfit(np.log(dts), np.log(errs), 1)[0]
    assert abs(slope - order) < 0.1


def test_match_endtime_lands_on_t_end():
    grid, B, U0 = _wet_basin(16, 12)
    o = oracle_2d(grid, system_2d(grid, B, sources=[VF.SourceTermBottom()]), RK2, U0)
    o.simulate_to_time(0.337)
    assert o.t == 0.337


# --------------------------------------------------------------- Float32
def test_float32_storage_tracks_float64():
    """make_cpu_backend(Float32): Float32 storage with the Float64 parameters
    (the reference's promotion rules); on a wet basin with bottom, rain and
    friction it stays within Float32 round-off of the Float64 run."""
    grid, B, U0 = _wet_basin()
    sysm = system_2d(grid, B, sources=[VF.SourceTermBottom(), VF.ConstantRain(0.036)],
                     friction=VF.ImplicitFriction(), kappa=1e-4)
    o64 = oracle_2d(grid, sysm, RK2, U0)
    o32 = oracle_2d(grid, sysm, RK2, U0, realtype=np.float32)
    o64.advance(50)
    o32.advance(50)
    assert o32.state().dtype == np.float32
    w64, w32 = o64.state()[2:-2, 2:-2, 0], o32.state()[2:-2, 2:-2, 0].astype(np.float64)
    assert np.max(np.abs(w64 - w32)) / np.max(np.abs(w64)) < 1e-5
    assert abs(o64.t - o32.t) / o64.t < 1e-5


# ---------------------------------------------------------- FunctionalRain
# test/test_rain.jl:98-160: the same lake, rain from a host function f(t, x, y)
# evaluated on every cell centre at the step's start time.
def _functional_rain_run(fn, hours, nx=10):
    grid = VF.CartesianGrid(nx, nx, gc=2, boundary=VF.WallBC(), extent=[[0.0, 1000.0], [0.0, 1000.0]])
    rain = VF.FunctionalRain(fn, grid)
    U0 = np.zeros((nx, nx, 3)); U0[:, :, 0] = 1.0
    o = oracle_2d(grid, system_2d(grid, sources=[rain]), FE, U0)
    T, t = hours * 3600.0, 0.0
    while t < T:
        o.set_rain_field(VF.rain_field_values(rain, grid, t))
        t += o.step(T - t, t)
    assert abs(t - T) <= 1e-10
    return o.state()[2:-2, 2:-2]


def test_functional_rain_constant():
    U = _functional_rain_run(lambda t, x, y: 0.01 / 3600.0, 3)
    assert np.max(np.abs(U[:, :, 0] - (3 * 0.01 + 1.0))) <= 1e-12
    assert np.all(U[:, :, 1] == 0.0) and np.all(U[:, :, 2] == 0.0)


def test_functional_rain_time_dependent():
    def fn(t, x, y):
        rate = 15.0 * 0.001 if t < 3600 else (0.3 * 0.001 if t < 7200 else 20.0 * 0.001)
        return rate / 3600
    U = _functional_rain_run(fn, 3)
    assert np.max(np.abs(U[:, :, 0] - (0.0353 + 1.0))) <= 1e-4


def test_functional_rain_spatial_total():
    def fn(t, x, y):
        return 3 * (x / 500) * (y / 500) * 0.01 / 3600.0 if 3600.0 < t < 7200.0 else 0.0 * x
    U = _functional_rain_run(fn, 3)
    assert abs(U[:, :, 0].sum() - 100 * 1.03) <= 1e-2


def test_functional_rain_equals_constant_rain_bitwise():
    """A FunctionalRain returning rate/3600 is ConstantRain(rate) bit for bit."""
    Uf = _functional_rain_run(lambda t, x, y: 0.036 / 3600, 1)
    Uc = _rain_run(VF.ConstantRain(0.036), 1)
    assert np.array_equal(Uf, Uc)


def test_cell_center_coordinates_follow_grid_jl():
    grid = VF.CartesianGrid(7, 5, gc=2, extent=[[0.5, 7.5], [-1.0, 9.0]])
    xs, ys = VF.model.cell_center_coordinates(grid)
    for i in range(grid.totalcells[0]):   # grid.jl:117-121 with 1-based I = i + 1
        assert xs[i] == 0.5 + grid.delta[0] * ((i + 1) - 0.5 - 2)
    for j in range(grid.totalcells[1]):
        assert ys[j] == -1.0 + grid.delta[1] * ((j + 1) - 0.5 - 2)
    assert np.allclose(xs[2:-2], grid.cell_centers()[:, 0, 0]) and np.allclose(ys[2:-2], grid.cell_centers()[0, :, 1])
    calls = []
    f = VF.rain_field_values(VF.FunctionalRain(lambda t, x, y: calls.append(1) or (x + 10 * y if x < 3 else 0.0), grid),
                             grid, 0.0)   # scalar-only function: evaluated per cell
    assert f.shape == grid.totalcells and len(calls) >= grid.totalcells[0] * grid.totalcells[1]
    assert f[0, 0] == xs[0] + 10 * ys[0]
