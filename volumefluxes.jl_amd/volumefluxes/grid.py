"""CartesianGrid and boundary-condition types (src/grid.jl).

Pure host-side description; the HIP backend reads nx, ny, gc, Δ and the
boundary kind from it.  Index conventions follow the reference but are
0-based: cell (i, j) with i in [0, nxT), x fastest.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


class BoundaryCondition:
    kind = -1


@dataclass(frozen=True)
class PeriodicBC(BoundaryCondition):   # grid.jl:21-22
    kind = 1


@dataclass(frozen=True)
class WallBC(BoundaryCondition):       # grid.jl:24-25
    kind = 0


@dataclass(frozen=True)
class NeumannBC(BoundaryCondition):    # grid.jl:27-28
    kind = 2


XDIR, YDIR = 0, 1


def _linrange(a: float, b: float, n: int) -> np.ndarray:
    """Julia's LinRange(a, b, n) element values: (1-t)*a + t*b with t=j/(n-1)."""
    d = n - 1
    out = np.empty(n)
    for j in range(n):
        t = j / d
        out[j] = (1 - t) * a + t * b
    return out


class CartesianGrid:
    """CartesianGrid(nx[, ny]; gc=1, boundary=PeriodicBC(), extent) (grid.jl:72-90)."""

    def __init__(self, nx: int, ny: int | None = None, *, gc: int = 1, boundary=None, extent=None):
        self.boundary = boundary if boundary is not None else PeriodicBC()
        self.gc = int(gc)
        if ny is None:
            ext = np.asarray(extent if extent is not None else [[0.0, 1.0]], dtype=float).reshape(-1)
            self.dimension = 1
            self.nx, self.ny = int(nx), 1
            self.extent = (float(ext[0]), float(ext[1]))
            self.delta = ((self.extent[1] - self.extent[0]) / nx,)
            self.totalcells = (self.nx + 2 * self.gc,)
        else:
            ext = np.asarray(extent if extent is not None else [[0.0, 1.0], [0.0, 1.0]], dtype=float)
            self.dimension = 2
            self.nx, self.ny = int(nx), int(ny)
            self.extent = (float(ext[0, 0]), float(ext[0, 1]), float(ext[1, 0]), float(ext[1, 1]))
            self.delta = ((self.extent[1] - self.extent[0]) / nx, (self.extent[3] - self.extent[2]) / ny)
            self.totalcells = (self.nx + 2 * self.gc, self.ny + 2 * self.gc)

    # grid.jl:63-69
    @property
    def size(self):
        return self.totalcells

    @property
    def ghostcells(self):
        return (self.gc,) * self.dimension

    def interior_size(self):
        return tuple(t - 2 * self.gc for t in self.totalcells)

    def compute_dx(self, direction: int = XDIR) -> float:
        return self.delta[direction]

    def compute_dy(self) -> float:
        return self.delta[YDIR]

    # grid.jl:103-115
    def cell_faces_dir(self, direction: int, interior: bool = True) -> np.ndarray:
        start = self.extent[2 * direction]
        end = self.extent[2 * direction + 1]
        if interior:
            return _linrange(start, end, self.interior_size()[direction] + 1)
        dx = self.compute_dx(direction)
        return _linrange(start - self.gc * dx, end + self.gc * dx, self.totalcells[direction] + 1)

    def cell_faces(self, interior: bool = True) -> np.ndarray:
        """2D: array (nfx, nfy, 2) of face corner coordinates (grid.jl:123-135).
        1D: vector of face positions (grid.jl:92-101)."""
        if self.dimension == 1:
            return self.cell_faces_dir(XDIR, interior)
        xf = self.cell_faces_dir(XDIR, interior)
        yf = self.cell_faces_dir(YDIR, interior)
        out = np.empty((len(xf), len(yf), 2))
        out[:, :, 0] = xf[:, None]
        out[:, :, 1] = yf[None, :]
        return out

    def cell_centers(self, interior: bool = True) -> np.ndarray:
        """2D: (n, m, 2) cell centres = face + (dx, dy)/2 (grid.jl:151-161).
        1D: faces[:-1] + (faces[1]-faces[0])/2.0 (grid.jl:139-143)."""
        if self.dimension == 1:
            f = self.cell_faces(interior)
            return f[:-1] + (f[1] - f[0]) / 2.0
        f = self.cell_faces(interior)
        c = f[:-1, :-1, :].copy()
        c[:, :, 0] = c[:, :, 0] + self.delta[0] / 2
        c[:, :, 1] = c[:, :, 1] + self.delta[1] / 2
        return c


def compute_dx(grid: CartesianGrid, direction: int = XDIR) -> float:
    return grid.compute_dx(direction)


def cell_centers(grid: CartesianGrid, interior: bool = True) -> np.ndarray:
    return grid.cell_centers(interior)


def cell_faces(grid: CartesianGrid, interior: bool = True) -> np.ndarray:
    return grid.cell_faces(interior)
