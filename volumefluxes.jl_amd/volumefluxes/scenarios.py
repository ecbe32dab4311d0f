"""Seeded synthetic inputs for the benchmark and test workloads.

terrain: bottom B on the (nxT+1, nyT+1) corners, ghosts included, = slope*x
plus n Gaussian bumps with amplitude U[0.5, 5] m, width U[20, 200] m and
centres uniform over the domain; initial surface w = max(B_cell, w0) with
hu = hv = 0, so cells with B_cell > w0 start dry.
dam_break: flat bottom, w = h_left left of x0 and h_right elsewhere.
terrain_workload: the single-grid terrain run on an n x n grid at 1 m
spacing (the configuration the throughput figures are quoted on).
serving_request / serving_requests: the flood scenarios of the serving
workload (dam breaks and terrains on square grids, Poisson arrivals).
"""
from __future__ import annotations

import numpy as np

from .grid import CartesianGrid


def corner_coordinates(grid: CartesianGrid):
    """x, y of every corner (i, j), ghosts included (cell_faces(grid, interior=false))."""
    xf = grid.cell_faces_dir(0, interior=False)
    yf = grid.cell_faces_dir(1, interior=False)
    return xf, yf


def cell_bottom(B: np.ndarray) -> np.ndarray:
    """B_cell on every cell, in the reference's summation order (bottom_topography.jl:97)."""
    return 0.25 * (((B[:-1, :-1] + B[1:, :-1]) + B[:-1, 1:]) + B[1:, 1:])


def terrain_topography(grid: CartesianGrid, *, seed: int = 0, n_bumps: int = 32, slope: float = 1e-3,
                       amp=(0.5, 5.0), width=(20.0, 200.0)) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x0, x1, y0, y1 = grid.extent
    A = rng.uniform(amp[0], amp[1], n_bumps)
    W = rng.uniform(width[0], width[1], n_bumps)
    cx = rng.uniform(x0, x1, n_bumps)
    cy = rng.uniform(y0, y1, n_bumps)
    xf, yf = corner_coordinates(grid)
    B = np.empty((len(xf), len(yf)), order="F")
    B[:, :] = slope * xf[:, None]
    for k in range(n_bumps):
        gx = np.exp(-((xf - cx[k]) ** 2) / (2.0 * W[k] ** 2))
        gy = np.exp(-((yf - cy[k]) ** 2) / (2.0 * W[k] ** 2))
        B += A[k] * gx[:, None] * gy[None, :]
    return B


def terrain_initial_state(grid: CartesianGrid, B: np.ndarray, w0: float = 2.0) -> np.ndarray:
    """Interior (nx, ny, 3): w = max(B_cell, w0), hu = hv = 0."""
    g = grid.gc
    Bc = cell_bottom(B)[g:-g, g:-g]
    U = np.zeros((grid.nx, grid.ny, 3), order="F")
    U[:, :, 0] = np.maximum(Bc, w0)
    return U


def dam_break_initial_state(grid: CartesianGrid, x0: float, h_left: float, h_right: float,
                            direction: int = 0) -> np.ndarray:
    c = grid.cell_centers()
    U = np.zeros((grid.nx, grid.ny, 3), order="F")
    U[:, :, 0] = np.where(c[:, :, direction] < x0, h_left, h_right)
    return U


def terrain_workload(n: int, precision: str = "f64", seed: int = 2048):
    """The single-grid terrain run: n x n cells at dx = dy = 1 m, wall
    boundaries, bottom = 1e-3 slope + 32 seeded bumps (amplitude U[0.5, 5] m,
    width U[20, 200] m), w = max(B_cell, 2 m), Manning n = 0.03 (bh2021 law,
    Cz = n^2), rain 1e-5 m/s (ConstantRain(1e-5 * 3600) m/h), kappa = 1e-4,
    depth cutoff 1e-5; RungeKutta2 at CFL 0.25 is the caller's stepper.
    Returns (grid, system, realtype, U0 interior, B corners)."""
    from . import model as M

    grid = CartesianGrid(n, n, gc=2, boundary=M.WallBC(), extent=[[0.0, float(n)], [0.0, float(n)]])
    B = terrain_topography(grid, seed=seed, n_bumps=32, slope=1e-3, amp=(0.5, 5.0), width=(20.0, 200.0))
    eq = M.ShallowWaterEquations(M.BottomTopography2D(B, None, grid), desingularizing_kappa=1e-4)
    manning_n = 0.03
    system = M.ConservedSystem(None, M.LinearReconstruction(), M.CentralUpwind(eq), eq, grid,
                               [M.SourceTermBottom(), M.ConstantRain(1e-5 * 3600)],
                               M.ImplicitFriction(Cz=manning_n * manning_n))
    realtype = np.float64 if precision == "f64" else np.float32
    return grid, system, realtype, terrain_initial_state(grid, B, 2.0), B


def serving_request(side: int, kind: str, topo_seed: int = 0):
    """One serving scenario on a side x side grid at dx = 1 m, wall boundaries,
    Manning n = 0.03 (bh2021), kappa = 1e-4, f64: kind "dam_break" (flat
    bottom, 2 m | 1 m at the middle, no sources) or "terrain" (seeded bumps
    as in terrain_workload, bottom source and rain 1e-5 m/s).
    Returns (grid, system, U0 interior)."""
    from . import model as M

    grid = CartesianGrid(side, side, gc=2, boundary=M.WallBC(), extent=[[0.0, float(side)], [0.0, float(side)]])
    friction = M.ImplicitFriction(Cz=0.03 * 0.03)
    if kind == "dam_break":
        eq = M.ShallowWaterEquations(desingularizing_kappa=1e-4)
        system = M.ConservedSystem(None, M.LinearReconstruction(), M.CentralUpwind(eq), eq, grid, [], friction)
        return grid, system, dam_break_initial_state(grid, side / 2.0, 2.0, 1.0)
    if kind != "terrain":
        raise ValueError(f"unknown serving scenario kind {kind!r}")
    B = terrain_topography(grid, seed=topo_seed, n_bumps=32, slope=1e-3, amp=(0.5, 5.0), width=(20.0, 200.0))
    eq = M.ShallowWaterEquations(M.BottomTopography2D(B, None, grid), desingularizing_kappa=1e-4)
    system = M.ConservedSystem(None, M.LinearReconstruction(), M.CentralUpwind(eq), eq, grid,
                               [M.SourceTermBottom(), M.ConstantRain(1e-5 * 3600)], friction)
    return grid, system, terrain_initial_state(grid, B, 2.0)


def serving_requests(rate: float, duration: float, t_range, seed: int, sides=(64, 2048)):
    """Poisson arrivals (`rate` per second over `duration` seconds); each a
    square grid of side log-uniform in `sides`, half dam breaks and half
    terrains, t_end uniform in `t_range`.  A list of dicts with arrival,
    grid, system, U0, kind, side and t_end."""
    rng = np.random.default_rng(seed)
    reqs, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > duration:
            break
        side = int(round(float(np.exp(rng.uniform(np.log(sides[0]), np.log(sides[1]))))))
        kind = "dam_break" if rng.uniform() < 0.5 else "terrain"
        topo_seed = int(rng.integers(1 << 30)) if kind == "terrain" else 0
        grid, system, U0 = serving_request(side, kind, topo_seed)
        reqs.append({"arrival": t, "grid": grid, "system": system, "U0": U0, "kind": kind, "side": side,
                     "t_end": float(rng.uniform(*t_range))})
    return reqs
