"""Terrain input and snapshot output, mirroring the reference's example
tools (examples/example_tools.jl): a delimited-text terrain loader, the
repeated 2x2 coarsening, per-cell bottom values, the total-water-volume
callback and per-time NPZ snapshots.  Host-side only."""
from __future__ import annotations

import math
import os

import numpy as np

from .model import BottomTopography2D


def load_grid(filename: str, delim: str = ",") -> np.ndarray:
    """loadgrid(filename; delim=',') (example_tools.jl:24-26): a delimited
    text matrix of Float64, rows as the file's lines."""
    return np.loadtxt(filename, delimiter=delim, dtype=np.float64, ndmin=2)


def coarsen(data: np.ndarray, times: int) -> np.ndarray:
    """coarsen(data, times) (example_tools.jl:29-45), as written: each pass
    halves both sizes (floor) and new cell (i, j) (1-based) sums old rows
    max(1, 2(i-1)) : min(n, 2(i-1)+1) and the same columns, each entry
    divided by 4.  The first row and column therefore gather one old row
    (column) only, the others two."""
    data = np.asarray(data, dtype=np.float64)
    for _ in range(int(times)):
        n0, n1 = data.shape
        m0, m1 = n0 // 2, n1 // 2
        out = np.zeros((m0, m1))
        for i in range(1, m0 + 1):
            rows = range(max(1, 2 * (i - 1)), min(n0, 2 * (i - 1) + 1) + 1)
            for j in range(1, m1 + 1):
                cols = range(max(1, 2 * (j - 1)), min(n1, 2 * (j - 1) + 1) + 1)
                acc = out[i - 1, j - 1]
                for r in rows:           # the reference's loop order: rows outer, columns inner
                    for c in cols:
                        acc += data[r - 1, c - 1] / 4
                out[i - 1, j - 1] = acc
        data = out
    return data


def bottom_per_cell(bottom) -> np.ndarray:
    """bottom_per_cell (example_tools.jl:54-80): B_cell of every cell of a
    corner array (or BottomTopography2D), ghosts included, in the reference's
    summation order (bottom_topography.jl:97)."""
    B = bottom.B if isinstance(bottom, BottomTopography2D) else np.asarray(bottom, dtype=np.float64)
    return 0.25 * (((B[:-1, :-1] + B[1:, :-1]) + B[:-1, 1:]) + B[1:, 1:])


class TotalWaterVolume:
    """TotalWaterVolume(bottom_topography) callback (example_tools.jl:48-52,
    81-100) without the plot: appends sum(h) over the interior (h = w -
    B_cell) and t after every call."""

    def __init__(self, bottom_topography):
        self.bottom_topography = bottom_topography
        self.water_volume: list[float] = []
        self.times: list[float] = []

    def __call__(self, t, sim):
        from .simulator import current_interior_state
        g = sim.grid.gc
        bottom = bottom_per_cell(self.bottom_topography)[g:-g, g:-g]
        h = current_interior_state(sim).h - bottom
        self.water_volume.append(float(np.sum(h)))
        self.times.append(float(t))


class SnapshotWriter:
    """The state callback of example_tools.jl:111-136 without the plot: at
    each call writes w_<T>.npz (the depth h = w - B_cell), hu_<T>.npz and
    hv_<T>.npz under directory/basename, T = ceil(t), each file one array in
    NPY format as the reference's npzwrite of a single array does.  Use it
    as IntervalWriter(step, SnapshotWriter(...))."""

    def __init__(self, directory: str, basename: str, bottom_topography):
        self.dir = os.path.join(directory, basename)
        self.bottom_topography = bottom_topography
        self.written: list[str] = []

    def __call__(self, t, sim):
        from .simulator import current_interior_state
        os.makedirs(self.dir, exist_ok=True)
        g = sim.grid.gc
        st = current_interior_state(sim)
        h = st.h - bottom_per_cell(self.bottom_topography)[g:-g, g:-g]
        tag = str(int(math.ceil(t)))
        for name, arr in (("w", h), ("hu", st.hu), ("hv", st.hv)):
            path = os.path.join(self.dir, f"{name}_{tag}.npz")
            with open(path, "wb") as f:
                np.save(f, np.asfortranarray(arr))
            self.written.append(path)
