"""Known-answer tests of the CPU oracle's scalar functions and index maps,
taken from the reference's own unit tests (file:line cited per test) or from
closed forms.  No GPU needed."""
import ctypes
import math

import numpy as np
import pytest

import oracle as O
import volumefluxes as VF
from volumefluxes import _capi as C


L = O.lib()


# ---- minmod: test/test_minmod.jl:24-27
def test_minmod_kats():
    assert L.oracle_minmod(1, 2, 3) == 1
    assert L.oracle_minmod(-1, -2, -3) == -1
    assert L.oracle_minmod(1, -2, 3) == 0
    got = [L.oracle_minmod(a, b, c) for a, b, c in zip([1, -2, -3], [3, -5, 2], [4, -3, 4])]
    assert got == [1, -2, 0]


def test_minmod_nan_and_zero():
    # any NaN makes every comparison false: the reference returns zero(a)
    assert L.oracle_minmod(math.nan, 1.0, 2.0) == 0.0
    r = L.oracle_minmod(0.0, 1.0, 2.0)
    assert r == 0.0 and not math.copysign(1, r) < 0


# ---- Julia max/min (IEEE 754-2019 maximum/minimum): NaN-propagating, -0 < +0
@pytest.mark.parametrize("a,b,mx,mn", [
    (1.0, 2.0, 2.0, 1.0),
    (-0.0, 0.0, 0.0, -0.0),
    (0.0, -0.0, 0.0, -0.0),
    (-3.0, -3.0, -3.0, -3.0),
    (math.inf, 1.0, math.inf, 1.0),
])
def test_jl_max_min(a, b, mx, mn):
    got_mx, got_mn = L.oracle_jl_max(a, b), L.oracle_jl_min(a, b)
    assert got_mx == mx and math.copysign(1, got_mx) == math.copysign(1, mx)
    assert got_mn == mn and math.copysign(1, got_mn) == math.copysign(1, mn)


def test_jl_max_min_nan():
    for a, b in [(math.nan, 1.0), (1.0, math.nan), (math.nan, -math.inf)]:
        assert math.isnan(L.oracle_jl_max(a, b))
        assert math.isnan(L.oracle_jl_min(a, b))


# ---- desingularize (equation.jl:56, Brodtkorb & Holm 2021 eq. 26): closed forms
@pytest.mark.parametrize("kappa", [1e-5, 1e-4, 0.3])
def test_desingularize_closed_forms(kappa):
    d = L.oracle_desingularize
    for h in [kappa, 2 * kappa, 1.0, 123.5]:
        assert d(h, kappa) == h            # |h| >= κ: h* = h
        assert d(-h, kappa) == -h
    assert d(0.0, kappa) == kappa / 2.0    # h = +0: κ/2
    assert d(-0.0, kappa) == -kappa / 2.0  # copysign(1, -0.0) = -1
    h = kappa / 3
    assert d(h, kappa) == h * h / (2 * kappa) + kappa / 2.0   # min(...) branch, < κ
    assert math.isnan(d(math.nan, kappa))
    # continuity at |h| = κ: h^2/(2κ) + κ/2 = κ exactly there
    assert d(np.nextafter(kappa, 0), kappa) <= kappa


# ---- rain: test/test_rain.jl:56-66
def _cfg_with(**kw):
    cfg = C.SfvConfig()
    cfg.nx, cfg.ny, cfg.gc = 4, 4, 2
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def test_constant_rain_kat():
    cfg = _cfg_with(rain=C.SFV_RAIN_CONSTANT, rain_rate=0.01)
    assert L.oracle_compute_rain(ctypes.byref(cfg), 12.6) == 0.01 / 3600.0
    assert VF.compute_rain(VF.ConstantRain(0.01), 12.6) == 0.01 / 3600.0


def test_time_dependent_rain_kats():
    rates = np.array([15.0, 0.3, 20.0]) * 0.001
    times = np.array([0.0, 1.0, 2.0]) * 3600
    cfg = _cfg_with(rain=C.SFV_RAIN_TIME_TABLE, rain_table_len=3, rain_table_rates=C.dptr(rates),
                    rain_table_times=C.dptr(times))
    want = {10: 0.015 / 3600.0, 3610: 0.0003 / 3600.0, 7210: 0.02 / 3600.0, 1.0e6: 0.02 / 3600.0,
            0.0: 0.015 / 3600.0}
    mirror = VF.TimeDependentRain(rates, times)
    for t, w in want.items():
        assert L.oracle_compute_rain(ctypes.byref(cfg), float(t)) == w
        assert VF.compute_rain(mirror, float(t)) == w


# ---- infiltration: test/test_infiltration.jl:58-61
def test_horton_kats():
    grid = VF.CartesianGrid(10, 10, gc=2, extent=[[0, 100], [0, 100]])
    inf = VF.HortonInfiltration(grid)
    assert np.all(inf.factor == 1.0) and inf.factor.shape == grid.size
    cfg = _cfg_with(infiltration=C.SFV_INFILTRATION_HORTON, horton_f0=inf.f0, horton_fc=inf.fc, horton_k=inf.k,
                    infiltration_factor_scalar=1.0)
    cfg.nx, cfg.ny = 10, 10
    assert L.oracle_compute_infiltration(ctypes.byref(cfg), 0.0, 0) == inf.f0
    assert L.oracle_compute_infiltration(ctypes.byref(cfg), 1e6, 0) == inf.fc


# ---- bottom lookups: test/test_bottom_topography.jl:87-113 (2D), 0-based indices
def test_bottom_topography_2d_kats():
    grid = VF.CartesianGrid(2, 2, gc=2, extent=[[0.0, 12.0], [0.0, 20.0]])
    faces = grid.cell_faces(interior=False)
    B = faces[:, :, 0] + faces[:, :, 1]
    cfg = _cfg_with()
    cfg.nx, cfg.ny = 2, 2
    o = O.Oracle2D(cfg, B)
    mirror = VF.BottomTopography2D(B, None, grid)
    tol = 1e-10
    cases = [  # (which, i, j, d, expected) from the 1-based KATs
        (0, 2, 2, 0, 8), (0, 2, 3, 0, 18), (0, 3, 2, 0, 14), (0, 3, 3, 0, 24),
        (1, 2, 2, 0, 5), (2, 2, 2, 0, 11), (1, 2, 2, 1, 3), (2, 2, 2, 1, 13),
        (1, 3, 3, 0, 21), (2, 3, 3, 0, 27), (1, 3, 3, 1, 19), (2, 3, 3, 1, 29),
    ]
    for which, i, j, d, want in cases:
        assert abs(o.B_lookup(which, i, j, d) - want) < tol
        m = [mirror.B_cell(i, j), mirror.B_face_left(i, j, d), mirror.B_face_right(i, j, d)][which]
        assert abs(m - want) < tol
    assert o.B_lookup(2, 2, 3, 0) == o.B_lookup(1, 3, 3, 0)
    assert o.B_lookup(2, 3, 2, 1) == o.B_lookup(1, 3, 3, 1)
    # collect_topography_cells: cell averages equal x + y at the centres
    cells = VF.collect_topography_cells(mirror, grid, interior=False)
    centres = grid.cell_centers(interior=False)
    assert np.allclose(cells, centres[:, :, 0] + centres[:, :, 1], atol=1e-14)


def test_bad_topography_size_is_domain_error():
    grid = VF.CartesianGrid(2, 2, gc=2)
    with pytest.raises(VF.DomainError):
        VF.BottomTopography2D(np.zeros((5, 5)), None, grid)


# ---- boundary conditions: test/test_update_bc.jl:79-117 (wall 2D) and
#      test/test_periodic_bc_2d.jl:26-74 (periodic, gc in {1, 2})
def _bc_oracle(nx, ny, gc, bc):
    cfg = _cfg_with(bc=bc, reconstruction=C.SFV_NO_RECONSTRUCTION)
    cfg.nx, cfg.ny, cfg.gc = nx, ny, gc
    return O.Oracle2D(cfg, np.zeros((nx + 2 * gc + 1, ny + 2 * gc + 1)))


def test_wall_bc_2d():
    nx, ny, gc = 10, 5, 2
    o = _bc_oracle(nx, ny, gc, C.SFV_BC_WALL)
    grid = VF.CartesianGrid(nx, ny, gc=gc, boundary=VF.WallBC())
    x = grid.cell_centers(interior=False)
    U = np.stack([x[:, :, 0], (x[:, :, 0] + x[:, :, 1]) * 10, x[:, :, 0] * (x[:, :, 1] - 5)], axis=-1)
    o.set_full(U)
    o.apply_bc()
    u = o.state()
    I = slice(2, -2)
    assert np.array_equal(u[I, I], U[I, I])
    # h copied, hu negated across x walls, hv negated across y walls (1-based KATs -> 0-based)
    for v, sx, sy in [(0, 1, 1), (1, -1, 1), (2, 1, -1)]:
        assert np.array_equal(u[1, I, v], sx * u[2, I, v])
        assert np.array_equal(u[0, I, v], sx * u[3, I, v])
        assert np.array_equal(u[nx + 3, I, v], sx * u[nx, I, v])
        assert np.array_equal(u[nx + 2, I, v], sx * u[nx + 1, I, v])
        assert np.array_equal(u[I, 1, v], sy * u[I, 2, v])
        assert np.array_equal(u[I, 0, v], sy * u[I, 3, v])
        assert np.array_equal(u[I, ny + 3, v], sy * u[I, ny, v])
        assert np.array_equal(u[I, ny + 2, v], sy * u[I, ny + 1, v])
    # corners are never visited
    assert np.array_equal(u[:2, :2], U[:2, :2]) and np.array_equal(u[-2:, -2:], U[-2:, -2:])


@pytest.mark.parametrize("gc", [1, 2])
def test_periodic_bc_2d(gc):
    nx, ny = 11, 8
    o = _bc_oracle(nx, ny, gc, C.SFV_BC_PERIODIC)
    data = -42 * np.ones((nx + 2 * gc, ny + 2 * gc))
    for j in range(gc, ny + gc):
        for i in range(gc, nx + gc):
            data[i, j] = (j + 1) * nx + (i + 1)
    U = np.stack([data, data, data], axis=-1)
    o.set_full(U)
    o.apply_bc()
    out = o.state()[:, :, 0]
    # test_periodic_bc_2d.jl:49-70 with 1-based n -> 0-based n-1
    for i in range(gc, nx + gc):
        for n in range(1, gc + 1):
            assert out[i, n - 1] == out[i, -(gc - n + 1) - gc]
            assert out[i, -(gc - n) - 1] == out[i, gc + n - 1]
    for i in range(gc, ny + gc):
        for n in range(1, gc + 1):
            assert out[n - 1, i] == out[-(gc - n + 1) - gc, i]
            assert out[-(gc - n) - 1, i] == out[gc + n - 1, i]


def test_neumann_bc_replays_reference_index_map():
    # bc.jl:49-60 as written: the low side mirrors the interior, the high side
    # copies ghost from ghost in order (g = 1 then 2), so both high ghosts end
    # up with the outermost ghost's pre-BC value.
    nx, ny, gc = 6, 4, 2
    o = _bc_oracle(nx, ny, gc, C.SFV_BC_NEUMANN)
    rng = np.random.default_rng(0)
    U = rng.standard_normal((nx + 4, ny + 4, 3))
    o.set_full(U)
    o.apply_bc()
    u = o.state()
    I = slice(2, -2)
    assert np.array_equal(u[0, I], U[3, I]) and np.array_equal(u[1, I], U[2, I])
    assert np.array_equal(u[nx + 2, I], U[nx + 3, I]) and np.array_equal(u[nx + 3, I], U[nx + 3, I])
    assert np.array_equal(u[I, 0], U[I, 3]) and np.array_equal(u[I, ny + 3], U[I, ny + 3])


# ---- central upwind: closed forms
def _cu(d, m, q, cutoff=1e-5, kappa=1e-5):
    cfg = _cfg_with(g=9.81, rho=1.0, depth_cutoff=cutoff, desingularizing_kappa=kappa)
    F = (ctypes.c_double * 3)()
    mm = (ctypes.c_double * 3)(*m)
    qq = (ctypes.c_double * 3)(*q)
    s = L.oracle_central_upwind(ctypes.byref(cfg), d, mm, qq, F)
    return np.array(F[:]), s


def test_central_upwind_equal_states_gives_physical_flux():
    h, hu, hv = 2.0, 0.0, 0.0
    F, s = _cu(0, (h, hu, hv), (h, hu, hv))
    c = math.sqrt(9.81 * h)
    assert s == c
    assert abs(F[1] - 0.5 * 9.81 * h * h) < 1e-12 and F[0] == 0.0 and F[2] == 0.0


def test_central_upwind_dry_faces():
    F, s = _cu(0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    assert s == 0.0 and np.all(F == 0.0)
    F, s = _cu(1, (1e-6, 0.0, 0.0), (-1e-7, 0.0, 0.0))
    assert s == 0.0 and np.all(F == 0.0)


def test_central_upwind_mirror_symmetry():
    m, q = (1.3, 0.4, -0.2), (0.7, -0.1, 0.3)
    F, s = _cu(0, m, q)
    # mirror x -> -x: swap sides and negate hu; mass and hv fluxes flip sign
    Fm, sm = _cu(0, (q[0], -q[1], q[2]), (m[0], -m[1], m[2]))
    assert s == sm
    assert np.allclose(Fm, [-F[0], F[1], -F[2]], rtol=1e-13, atol=1e-15)
    # x and y faces are the same computation with hu and hv exchanged
    Fy, sy = _cu(1, (m[0], m[2], m[1]), (q[0], q[2], q[1]))
    assert sy == s and np.array_equal(Fy, [F[0], F[2], F[1]])
