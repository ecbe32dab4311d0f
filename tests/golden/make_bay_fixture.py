"""Writes tests/golden/bay_terrain.npz: the terrain of the reference's bay
example (examples/terrain.jl:59-74) as that example builds it, for the GPU
test of that configuration.

Steps, each as the example does them:
  1. loadgrid(".../examples/data/bay.txt") (example_tools.jl:24-26): a
     comma-separated 451 x 601 matrix of elevations;
  2. coarsen(terrain, 2) (example_tools.jl:29-45): two 2x2 passes, giving
     112 x 150;
  3. naive_infill (terrain.jl:37-56): every entry below 1e-3 becomes
     -log(d), d the Euclidean index distance to the nearest entry >= 1e-3
     (Julia's log and numpy's may differ in the last bit; the array is an
     input of the test, given to the HIP backend and the oracle alike).

The saved arrays: `terrain` (the corner array the example passes to
BottomTopography2D) and `upper_corner` (the original size, which the example
uses as the grid's extent).  Run it from the repository root in the build
container, where the reference checkout is at /root/reference.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "volumefluxes.jl_amd"))

from volumefluxes.io import coarsen, load_grid  # noqa: E402

SOURCE = "/root/reference/examples/data/bay.txt"


def naive_infill(terrain: np.ndarray) -> np.ndarray:
    out = terrain.copy()
    low = np.argwhere(terrain < 1e-3)
    high = np.argwhere(terrain >= 1e-3).astype(np.int64)
    for a in range(0, len(low), 256):
        blk = low[a:a + 256].astype(np.int64)
        d2 = ((blk[:, None, 0] - high[None, :, 0]) ** 2 + (blk[:, None, 1] - high[None, :, 1]) ** 2).min(axis=1)
        out[blk[:, 0], blk[:, 1]] = -np.log(np.sqrt(d2.astype(np.float64)))
    return out


def main() -> None:
    raw = load_grid(SOURCE)
    upper = np.array(raw.shape, dtype=np.float64)
    terrain = naive_infill(coarsen(raw, 2))
    path = os.path.join(HERE, "bay_terrain.npz")
    np.savez_compressed(path, terrain=np.asfortranarray(terrain), upper_corner=upper)
    print(f"{path}: terrain {terrain.shape}, min {terrain.min():.4f}, max {terrain.max():.4f}, "
          f"{int(np.sum(raw < 1e-3))} low entries before coarsening")


if __name__ == "__main__":
    main()
