"""Declarative model objects mirroring the reference's constructors.

Bottom topography (src/bottom_topography.jl), ShallowWaterEquations
(src/equations/swe_2D.jl), reconstructions (src/reconstruction/),
CentralUpwind (src/numericalflux/swe/centralupwind.jl), source terms
(src/sourceterms/), ImplicitFriction (src/friction.jl), ConservedSystem
(src/system.jl) and the time steppers (src/timestepper/).  These objects only
describe a run; :func:`build_config` turns them into the C ABI's sfv_config.
Validation raises the reference's exception types at the same points.
"""
from __future__ import annotations

import numpy as np

from . import _capi as C
from .grid import CartesianGrid, NeumannBC, PeriodicBC, WallBC

DomainError = C.DomainError
ArgumentError = C.ArgumentError


def _domain_error(msg):
    return C.DomainError(C.SFV_ERR_DOMAIN, msg)


def _argument_error(msg):
    return C.ArgumentError(C.SFV_ERR_ARGUMENT, msg)


# ------------------------------------------------------------- topography
class ConstantBottomTopography:
    """ConstantBottomTopography(B=0.0) (bottom_topography.jl:24-27)."""

    def __init__(self, B: float = 0.0):
        self.B = float(B)

    def is_zero(self) -> bool:
        return self.B == 0.0

    def corners(self, grid: CartesianGrid) -> np.ndarray:
        return np.full(tuple(t + 1 for t in grid.totalcells), self.B, order="F")


def validate_bottom_topography(B: np.ndarray, grid: CartesianGrid) -> None:
    """bottom_topography.jl:78-82."""
    want = tuple(t + 1 for t in grid.totalcells)
    if tuple(B.shape) != want:
        raise _domain_error(f"Bottom topography should be of size {want} but got {tuple(B.shape)}")


class BottomTopography1D:
    """BottomTopography1D(B, backend, grid) (bottom_topography.jl:31-37): B
    on the nxT+1 cell faces, ghosts included."""

    def __init__(self, B, backend=None, grid: CartesianGrid | None = None):
        B = np.asarray(B, dtype=np.float64)
        if grid is not None:
            validate_bottom_topography(B, grid)
        self.B = B.copy()

    def is_zero(self) -> bool:
        return bool(np.all(self.B == 0.0))

    def corners(self, grid: CartesianGrid) -> np.ndarray:
        validate_bottom_topography(self.B, grid)
        return self.B


class BottomTopography2D:
    """BottomTopography2D(B, backend, grid) (bottom_topography.jl:51-57): B is
    the corner array (nxT+1, nyT+1), indexed B[i, j] with x first."""

    def __init__(self, B, backend=None, grid: CartesianGrid | None = None):
        B = np.asarray(B, dtype=np.float64)
        if grid is not None:
            validate_bottom_topography(B, grid)
        self.B = np.asfortranarray(B)

    def is_zero(self) -> bool:
        return bool(np.all(self.B == 0.0))

    def corners(self, grid: CartesianGrid) -> np.ndarray:
        validate_bottom_topography(self.B, grid)
        return self.B

    # bottom_topography.jl:97-101 (0-based cell indices)
    def B_cell(self, i: int, j: int) -> float:
        B = self.B
        return 0.25 * (B[i, j] + B[i + 1, j] + B[i, j + 1] + B[i + 1, j + 1])

    def B_face_left(self, i: int, j: int, direction: int) -> float:
        B = self.B
        return 0.5 * (B[i, j] + B[i, j + 1]) if direction == 0 else 0.5 * (B[i, j] + B[i + 1, j])

    def B_face_right(self, i: int, j: int, direction: int) -> float:
        B = self.B
        return 0.5 * (B[i + 1, j] + B[i + 1, j + 1]) if direction == 0 else 0.5 * (B[i, j + 1] + B[i + 1, j + 1])


def collect_topography_cells(B, grid: CartesianGrid, interior: bool = True) -> np.ndarray:
    """bottom_topography.jl:137-146: average along x, then along y."""
    Bc = B.corners(grid)
    topo_y = 0.5 * (Bc[:-1, :] + Bc[1:, :])
    topo = 0.5 * (topo_y[:, :-1] + topo_y[:, 1:])
    if interior:
        g = grid.gc
        return topo[g:-g, g:-g]
    return topo


# --------------------------------------------------------------- equations
class ShallowWaterEquations:
    """ShallowWaterEquations(B; ρ=1.0, g=9.81, depth_cutoff=10^-5,
    desingularizing_kappa=10^-5) (swe_2D.jl:22-29).  Julia's 10^-5 is the
    double 1e-5."""

    def __init__(self, B=None, *, rho: float = 1.0, g: float = 9.81, depth_cutoff: float = 1e-5,
                 desingularizing_kappa: float = 1e-5):
        self.B = B if B is not None else ConstantBottomTopography()
        self.rho = float(rho)
        self.g = float(g)
        self.depth_cutoff = float(depth_cutoff)
        self.desingularizing_kappa = float(desingularizing_kappa)


class ShallowWaterEquations1D(ShallowWaterEquations):
    """ShallowWaterEquations1D(B; ρ, g, depth_cutoff, desingularizing_kappa)
    (swe_1D.jl:21-29): state (w, hu) on a CartesianGrid{1}."""


class LinearReconstruction:
    """LinearReconstruction(theta=1.2) (linearreconstruction.jl:22-25)."""

    def __init__(self, theta: float = 1.2):
        self.theta = float(theta)


class NoReconstruction:
    """NoReconstruction() (noreconstruction.jl:21)."""


class CentralUpwind:
    """CentralUpwind(eq) (centralupwind.jl:21-23)."""

    def __init__(self, eq):
        self.eq = eq


# ------------------------------------------------------------ source terms
class SourceTermBottom:
    """Well-balanced bottom slope source (source_term_bottom.jl:24-41)."""


class ConstantRain:
    """ConstantRain(rate): adds rate/3600 to w every step (source_term_rain.jl:24-30)."""

    def __init__(self, rain_rate: float):
        self.rain_rate = float(rain_rate)


def compute_rain(rain, t: float = 0.0, *index) -> float:
    """source_term_rain.jl:30 and :45-54."""
    if isinstance(rain, ConstantRain):
        return rain.rain_rate / 3600
    index = 0
    while index + 1 < len(rain.time):
        if t < rain.time[index + 1]:
            break
        index += 1
    return rain.rain_rates[index] / 3600


class TimeDependentRain:
    """TimeDependentRain(rates, time=[0.0]) step table (source_term_rain.jl:32-41)."""

    def __init__(self, rain_rates, time=(0.0,)):
        rates = np.asarray(rain_rates, dtype=np.float64).reshape(-1)
        times = np.asarray(time, dtype=np.float64).reshape(-1)
        if np.shape(rain_rates) != np.shape(time):
            raise _domain_error("Dimensions of input parameters rain_rates and time does not match")
        self.rain_rates = np.ascontiguousarray(rates)
        self.time = np.ascontiguousarray(times)


class FunctionalRain:
    """FunctionalRain(rain_function, grid) (source_term_rain.jl:57-68): w gains
    rain_function(t, x, y) on every cell, ghosts included, at the cell centre
    (cell_center, grid.jl:117-121), with t the step's start time.  The HIP
    backend evaluates the host function before each step and hands the
    rates to the kernels as a per-cell field (rain_field_values)."""

    def __init__(self, rain_function, grid):
        self.rain_function = rain_function
        self.grid = grid


def cell_center_coordinates(grid: CartesianGrid):
    """x, y of cell_center(grid, I) for every cell, ghosts included
    (grid.jl:117-121): start + Δ*(I - 0.5 - gc) with 1-based I, the
    half-integer formed exactly before the product."""
    x0, _, y0, _ = grid.extent
    dx, dy = grid.delta
    i1 = np.arange(grid.totalcells[0], dtype=np.float64) + 1.0
    j1 = np.arange(grid.totalcells[1], dtype=np.float64) + 1.0
    xs = x0 + dx * ((i1 - 0.5) - grid.gc)
    ys = y0 + dy * ((j1 - 0.5) - grid.gc)
    return xs, ys


_RAIN_POOL = None


def _rain_pool(workers: int):
    global _RAIN_POOL
    if _RAIN_POOL is None or _RAIN_POOL._max_workers != workers:
        from concurrent.futures import ThreadPoolExecutor
        _RAIN_POOL = ThreadPoolExecutor(max_workers=workers)
    return _RAIN_POOL


def rain_field_values(rain: FunctionalRain, grid: CartesianGrid, t: float, out=None, workers=None) -> np.ndarray:
    """compute_rain(rain, t, I) (source_term_rain.jl:64-68) for every cell, as
    a (nxT, nyT) float64 array (written into `out` when given, e.g. a pinned
    host buffer).  The function is called with array arguments when it
    accepts them, on blocks of columns evaluated by a pool of host threads
    (numpy releases the interpreter lock in its element-wise kernels; the
    rate of a cell depends on (t, x, y) alone, as the reference evaluates it
    per cell), else once per cell."""
    import os
    xs, ys = cell_center_coordinates(grid)
    nxT, nyT = grid.totalcells
    if out is None:
        out = np.empty((nxT, nyT), order="F")
    if workers is None:
        workers = max(1, min(16, len(os.sched_getaffinity(0))))

    def block(j0, j1):
        X = np.broadcast_to(xs[:, None], (nxT, j1 - j0))
        Y = np.broadcast_to(ys[None, j0:j1], (nxT, j1 - j0))
        v = np.asarray(rain.rain_function(t, X, Y), dtype=np.float64)
        if v.shape == ():
            out[:, j0:j1] = float(v)
        elif v.shape == (nxT, j1 - j0):
            out[:, j0:j1] = v
        else:
            return False
        return True

    try:
        if workers <= 1 or nxT * nyT < (1 << 16):
            ok = block(0, nyT)
        else:
            n = min(nyT, 4 * workers)
            edges = [nyT * k // n for k in range(n + 1)]
            ok = all(_rain_pool(workers).map(lambda k: block(edges[k], edges[k + 1]), range(n)))
        if ok:
            return out
    except (TypeError, ValueError):
        pass
    f = rain.rain_function
    for j in range(nyT):
        for i in range(nxT):
            out[i, j] = float(f(t, float(xs[i]), float(ys[j])))
    return out


class HortonInfiltration:
    """HortonInfiltration(grid, backend; factor=1.0, fc, f0, k)
    (source_term_infiltration.jl:21-47): rate factor*(fc + (f0-fc)*exp(-k t))."""

    def __init__(self, grid: CartesianGrid, backend=None, *, factor=1.0, fc: float = 3.272e-5,
                 f0: float = 1.977e-4, k: float = 2.43e-3):
        self.f0, self.fc, self.k = float(f0), float(fc), float(k)
        if np.ndim(factor) == 0:
            self.factor = np.full(grid.totalcells, float(factor), order="F")
            self.scalar_factor = float(factor)
        else:
            factor = np.asarray(factor, dtype=np.float64)
            if tuple(factor.shape) != tuple(grid.totalcells):
                raise _domain_error(
                    f"Infiltration factor should be of size {tuple(grid.totalcells)} but got {tuple(factor.shape)}")
            self.factor = np.asfortranarray(factor)
            self.scalar_factor = None


def compute_infiltration(f, t: float, index) -> float:
    """source_term_infiltration.jl:68-81 (index is a 0-based (i, j))."""
    if isinstance(f, HortonInfiltration):
        return f.factor[index] * (f.fc + (f.f0 - f.fc) * np.exp(-f.k * t))
    return f.infiltration_rate


class ConstantInfiltration:
    """ConstantInfiltration(rate) (source_term_infiltration.jl:73-81)."""

    def __init__(self, infiltration_rate: float):
        self.infiltration_rate = float(infiltration_rate)


# ------------------------------------------------------------------ friction
def friction_bh2021(c, h, speed):
    """friction.jl:40-43."""
    return -c * speed / (h * h)


def friction_bsa2012(c, h, speed):
    """friction.jl:30-33."""
    return -c * speed / (np.cbrt(h) * h)


def friction_fcg2016(c, h, speed):
    """friction.jl:35-38."""
    return -c * speed / (np.cbrt(h) * h * h)


_FRICTION_CODES = {friction_bh2021: C.SFV_FRICTION_BH2021, friction_bsa2012: C.SFV_FRICTION_BSA2012,
                   friction_fcg2016: C.SFV_FRICTION_FCG2016}


class ImplicitFriction:
    """ImplicitFriction(; Cz=0.03^2, friction_function=friction_bh2021) (friction.jl:23-27)."""

    def __init__(self, *, Cz: float = 0.03 * 0.03, friction_function=friction_bh2021):
        if friction_function not in _FRICTION_CODES:
            raise _argument_error("friction_function must be friction_bh2021, friction_bsa2012 or friction_fcg2016")
        self.Cz = float(Cz)
        self.friction_function = friction_function


# ------------------------------------------------------------------ steppers
class RungeKutta2:
    """RungeKutta2() (rungekutta.jl:21-24), two substeps."""
    substeps = 2


class ForwardEulerStepper:
    """ForwardEulerStepper() (forwardeuler.jl:23-26), one substep."""
    substeps = 1


def number_of_substeps(stepper) -> int:
    return stepper.substeps


# -------------------------------------------------------------------- system
class ConservedSystem:
    """ConservedSystem(backend, rec, flux, eq, grid, sources=[], implicit_source=None)
    (system.jl:24-54).  Checks as the reference does: B's size against the
    grid (is_compatible(eq, grid)) and a non-zero B needs SourceTermBottom
    (equation.jl:74-84)."""

    def __init__(self, backend, reconstruction, numericalflux, equation, grid, source_terms=(),
                 implicit_source_term=None):
        if not isinstance(source_terms, (list, tuple)):
            source_terms = [source_terms]
        self.backend = backend
        self.reconstruction = reconstruction
        self.numericalflux = numericalflux
        self.equation = equation
        self.grid = grid
        self.source_terms = list(source_terms)
        self.implicit_source_term = implicit_source_term
        if isinstance(equation.B, (BottomTopography1D, BottomTopography2D)):
            validate_bottom_topography(equation.B.B, grid)
        if not equation.B.is_zero() and not any(isinstance(s, SourceTermBottom) for s in self.source_terms):
            raise _argument_error(
                "Found non-zero bottom topography in equation, but no corresponding source term. "
                "Did you forget to add a source term to your system?")


# ------------------------------------------------------------------ 1D runs
EMBED_ROWS = 3


def embed_1d(system: ConservedSystem, grid: CartesianGrid):
    """The 2D system on an (nx, 3) grid that reproduces a CartesianGrid{1}
    system bitwise in its middle row: the reference's own check is that the
    2D scheme on (n, 3) with WallBC equals the 1D scheme bit for bit, dt
    included (test_swashes.jl:317-328).  Every row carries the 1D state (hv =
    0) and the 1D bottom, so all y flux differences vanish exactly; dy is
    dx * 2^60, so the y wave speeds never set dt (min(dx/a_x, dy/a_y) =
    dx/a_x, as in 1D).  The reference's identity holds for a flat bottom
    only: the 2D cell bottom 0.25*(((a+b)+c)+d) rounds differently from the
    1D 0.5*(a+b) (bottom_topography.jl:92,97), so the embedded system is
    marked and B_cell takes the 1D lookup.  Returns (grid2d, system2d)."""
    if grid.dimension != 1:
        raise _argument_error("embed_1d takes a CartesianGrid{1}")
    x0, x1 = grid.extent
    dy = grid.delta[0] * 2.0 ** 60
    g2 = CartesianGrid(grid.nx, EMBED_ROWS, gc=grid.gc, boundary=grid.boundary,
                       extent=[[x0, x1], [0.0, EMBED_ROWS * dy]])
    eq = system.equation
    nyc = g2.totalcells[1] + 1
    if isinstance(eq.B, ConstantBottomTopography):
        topo = ConstantBottomTopography(eq.B.B)
    else:
        B1 = eq.B.corners(grid)
        topo = BottomTopography2D(np.asfortranarray(np.repeat(B1[:, None], nyc, axis=1)), None, g2)
    eq2 = ShallowWaterEquations(topo, rho=eq.rho, g=eq.g, depth_cutoff=eq.depth_cutoff,
                                desingularizing_kappa=eq.desingularizing_kappa)
    sources = []
    for st in system.source_terms:
        if isinstance(st, HortonInfiltration):
            fac = np.asfortranarray(np.repeat(st.factor.reshape(-1)[:, None], g2.totalcells[1], axis=1))
            sources.append(HortonInfiltration(g2, factor=fac, fc=st.fc, f0=st.f0, k=st.k))
        elif isinstance(st, FunctionalRain):
            raise _argument_error("FunctionalRain takes (t, x, y); a 1D run has no y")
        else:
            sources.append(st)
    sys2 = ConservedSystem(None, system.reconstruction, CentralUpwind(eq2), eq2, g2, sources,
                           system.implicit_source_term)
    sys2.embedded_1d = True           # B_cell takes the 1D lookup (build_config)
    return g2, sys2


def build_config(system: ConservedSystem, stepper, grid: CartesianGrid, cfl: float, precision: int) -> C.SfvConfig:
    """Translate a reference-style system description into sfv_config."""
    if grid.dimension != 2:
        raise _argument_error("the HIP backend runs CartesianGrid{2}")
    eq = system.equation
    cfg = C.SfvConfig()
    cfg.nx, cfg.ny, cfg.gc = grid.nx, grid.ny, grid.gc
    cfg.dx, cfg.dy = grid.delta
    cfg.bc = grid.boundary.kind
    cfg.g, cfg.rho = eq.g, eq.rho
    cfg.depth_cutoff, cfg.desingularizing_kappa = eq.depth_cutoff, eq.desingularizing_kappa
    if isinstance(system.reconstruction, LinearReconstruction):
        cfg.reconstruction, cfg.theta = C.SFV_LINEAR_RECONSTRUCTION, system.reconstruction.theta
    else:
        cfg.reconstruction, cfg.theta = C.SFV_NO_RECONSTRUCTION, 0.0
    cfg.stepper = C.SFV_RUNGE_KUTTA2 if isinstance(stepper, RungeKutta2) else C.SFV_FORWARD_EULER
    cfg.cfl = float(cfl)
    cfg.bottom_source = int(any(isinstance(s, SourceTermBottom) for s in system.source_terms))
    fr = system.implicit_source_term
    if fr is None:
        cfg.friction, cfg.Cz = C.SFV_FRICTION_NONE, 0.0
    else:
        cfg.friction, cfg.Cz = _FRICTION_CODES[fr.friction_function], fr.Cz
    rains = [s for s in system.source_terms if isinstance(s, (ConstantRain, TimeDependentRain, FunctionalRain))]
    infs = [s for s in system.source_terms if isinstance(s, (HortonInfiltration, ConstantInfiltration))]
    if len(rains) > 1 or len(infs) > 1:
        raise _argument_error("the HIP backend takes at most one rain and one infiltration source")
    keep = []
    if rains:
        r = rains[0]
        if isinstance(r, FunctionalRain):
            cfg.rain = C.SFV_RAIN_FIELD
        elif isinstance(r, ConstantRain):
            cfg.rain, cfg.rain_rate = C.SFV_RAIN_CONSTANT, r.rain_rate
        else:
            cfg.rain = C.SFV_RAIN_TIME_TABLE
            cfg.rain_table_len = len(r.rain_rates)
            cfg.rain_table_rates, cfg.rain_table_times = C.dptr(r.rain_rates), C.dptr(r.time)
            keep += [r.rain_rates, r.time]
    cfg.horton_f0, cfg.horton_fc, cfg.horton_k = 1.977e-4, 3.272e-5, 2.43e-3
    cfg.infiltration_factor_scalar = 1.0
    if infs:
        f = infs[0]
        if isinstance(f, HortonInfiltration):
            cfg.infiltration = C.SFV_INFILTRATION_HORTON
            cfg.horton_f0, cfg.horton_fc, cfg.horton_k = f.f0, f.fc, f.k
            fac = np.ascontiguousarray(f.factor.reshape(-1, order="F"))
            cfg.infiltration_factor = C.dptr(fac)
            keep.append(fac)
        else:
            cfg.infiltration, cfg.infiltration_rate = C.SFV_INFILTRATION_CONSTANT, f.infiltration_rate
    order = [s for s in system.source_terms if s in rains or s in infs]
    cfg.rain_before_infiltration = 1 if (not order or not infs or order[0] in rains) else 0
    cfg.precision = precision
    cfg.embedded_1d = 1 if getattr(system, "embedded_1d", False) else 0
    cfg.keepalive(*keep)
    return cfg
