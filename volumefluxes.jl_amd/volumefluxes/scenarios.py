"""Seeded synthetic inputs for the benchmark and test workloads.

terrain: bottom B on the (nxT+1, nyT+1) corners, ghosts included, = slope*x
plus n Gaussian bumps with amplitude U[0.5, 5] m, width U[20, 200] m and
centres uniform over the domain; initial surface w = max(B_cell, w0) with
hu = hv = 0, so cells with B_cell > w0 start dry.
dam_break: flat bottom, w = h_left left of x0 and h_right elsewhere.
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
