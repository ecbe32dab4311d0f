"""Example I/O tools (examples/example_tools.jl): loader, coarsen, per-cell
bottom, total-water callback and NPZ snapshots.  CPU only."""
import numpy as np

import volumefluxes as VF
from volumefluxes.scenarios import cell_bottom


def test_load_grid_roundtrip(tmp_path):
    A = np.arange(12.0).reshape(3, 4) / 7.0
    p = tmp_path / "bay.txt"
    np.savetxt(p, A, delimiter=",", fmt="%.17g")
    assert np.array_equal(VF.load_grid(str(p)), A)
    np.savetxt(tmp_path / "b.txt", A, delimiter=" ", fmt="%.17g")
    assert np.array_equal(VF.load_grid(str(tmp_path / "b.txt"), delim=" "), A)


def test_coarsen_follows_the_reference_indexing():
    # example_tools.jl:34-40: new (i, j) gathers old rows max(1,2(i-1)):min(n,2(i-1)+1)
    d = np.arange(1.0, 17.0).reshape(4, 4)           # d[r-1, c-1] = 4(r-1) + c
    c = VF.coarsen(d, 1)
    assert c.shape == (2, 2)
    assert c[0, 0] == d[0, 0] / 4                     # row 1, column 1 only
    assert c[0, 1] == d[0, 1] / 4 + d[0, 2] / 4       # row 1, columns 2..3
    assert c[1, 0] == d[1, 0] / 4 + d[2, 0] / 4       # rows 2..3, column 1
    assert c[1, 1] == ((d[1, 1] / 4 + d[1, 2] / 4) + d[2, 1] / 4) + d[2, 2] / 4
    odd = np.ones((5, 3))
    c2 = VF.coarsen(odd, 1)
    assert c2.shape == (2, 1) and c2[0, 0] == 0.25 and c2[1, 0] == 0.5
    assert VF.coarsen(np.ones((16, 16)), 2).shape == (4, 4)


def test_bottom_per_cell_is_the_cell_lookup():
    B = np.random.default_rng(0).normal(size=(9, 7))
    assert np.array_equal(VF.bottom_per_cell(B), cell_bottom(B))
    grid = VF.CartesianGrid(4, 2, gc=2)
    assert np.array_equal(VF.bottom_per_cell(VF.BottomTopography2D(B, None, grid)), cell_bottom(B))


class _StubSim:
    """Just enough of a Simulator for the callbacks: grid and full state."""

    def __init__(self, grid, U):
        self.grid, self._U = grid, U

    def get_full_state(self):
        return self._U


def test_total_water_volume_and_snapshots(tmp_path):
    grid = VF.CartesianGrid(6, 5, gc=2, extent=[[0.0, 6.0], [0.0, 5.0]])
    B = np.fromfunction(lambda i, j: 0.1 * i + 0.05 * j, (11, 10))
    U = np.zeros((10, 9, 3), order="F")
    U[:, :, 0] = 2.0
    U[:, :, 1] = 0.5
    sim = _StubSim(grid, U)
    twv = VF.TotalWaterVolume(B)
    twv(1.5, sim)
    h = 2.0 - cell_bottom(B)[2:-2, 2:-2]
    assert twv.times == [1.5] and twv.water_volume == [float(np.sum(h))]
    w = VF.SnapshotWriter(str(tmp_path), "run", B)
    w(2.2, sim)
    got = np.load(tmp_path / "run" / "w_3.npz")
    assert got.shape == (6, 5) and np.array_equal(got, h)
    assert np.all(np.load(tmp_path / "run" / "hu_3.npz") == 0.5)
    assert len(w.written) == 3
