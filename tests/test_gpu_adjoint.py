"""The GPU adjoint against the oracle's forward-mode derivative
(oracle/oracle_ad.hpp): for random directions (δB, δCz, δU0),
<dJ/dB, δB> + dJ/dCz δCz + <dJ/dU0, δU0> must equal the directional
derivative the oracle carries through the same run.  Both are exact
derivatives of the same branch decisions, so they agree to rounding."""
import numpy as np
import pytest

import oracle as O
import volumefluxes as VF
from volumefluxes.scenarios import cell_bottom

from helpers import dam_break, make_hip, terrain
from test_oracle_ad import smooth_flow

pytestmark = pytest.mark.gpu


def _target(sc):
    g = sc.grid.gc
    return sc.U0[:, :, 0] - cell_bottom(sc.corners())[g:-g, g:-g] - 0.3


def _check(sc, nsteps, ce, tol=1e-9, dirs=3):
    tgt = _target(sc)
    sim = make_hip(sc)
    res = VF.adjoint_gradient(sim, nsteps, tgt, checkpoint_every=ce)
    cfg = sc.config()
    cfg._k = cfg
    J0, _, U_T, _ = O.ad_run(cfg, sc.corners(), sc.U0, nsteps, tgt)
    assert abs(res["J"] - J0) <= 1e-12 * abs(J0)
    assert np.array_equal(sim.get_full_state(), U_T)          # the handle ends at the final state
    rng = np.random.default_rng(11)
    B = sc.corners()
    for k in range(dirs):
        dB = rng.normal(size=B.shape)
        dU = rng.normal(size=sc.U0.shape) * 1e-2
        dC = float(rng.normal()) * 1e-3
        _, dJ, _, _ = O.ad_run(cfg, B, sc.U0, nsteps, tgt, dB=dB, dCz=dC, dU0=dU)
        got = float(np.sum(res["dJ_dB"] * dB) + res["dJ_dCz"] * dC + np.sum(res["dJ_dU0"] * dU))
        print(sc.name, k, got, dJ)
        assert abs(got - dJ) <= tol * max(abs(dJ), 1e-300), (got, dJ)
    return res


def test_adjoint_smooth_wall():
    _check(smooth_flow(terrain(40, 30, seed=21)), 17, 5)


def test_adjoint_smooth_periodic():
    _check(smooth_flow(terrain(36, 28, seed=22, bc=VF.PeriodicBC())), 12, 4)


def test_adjoint_wet_dry_terrain():
    """Dry bumps, rain, friction: wet/dry fronts, clamps and cut-offs."""
    _check(terrain(48, 40, seed=23), 20, 6, tol=1e-8)


def test_adjoint_dam_break_wet_dry():
    _check(dam_break(60, 12, hl=0.005, hr=0.0), 25, 10, tol=1e-8)


def test_adjoint_single_corners_and_friction():
    sc = smooth_flow(terrain(30, 26, seed=24))
    tgt = _target(sc)
    res = VF.adjoint_gradient(make_hip(sc), 10, tgt, checkpoint_every=3)
    cfg = sc.config()
    cfg._k = cfg
    B = sc.corners()
    for (i, j) in [(0, 0), (2, 2), (5, 7), (17, 13), (33, 29)]:
        e = np.zeros_like(B)
        e[i, j] = 1.0
        _, dJ, _, _ = O.ad_run(cfg, B, sc.U0, 10, tgt, dB=e)
        assert abs(res["dJ_dB"][i, j] - dJ) <= 1e-9 * abs(dJ) + 1e-15, (i, j, res["dJ_dB"][i, j], dJ)
    _, dJc, _, _ = O.ad_run(cfg, B, sc.U0, 10, tgt, dCz=1.0)
    assert abs(res["dJ_dCz"] - dJc) <= 1e-9 * abs(dJc)
    assert VF.manning_gradient(res["dJ_dCz"], 0.03) == 2 * 0.03 * res["dJ_dCz"]


def test_adjoint_recompute_path_equals_kept_path(monkeypatch):
    """Segments either keep their stage-1 states and right-hand sides from the
    recompute (on a second stream, overlapped with the previous segment's
    reverse steps, or on one stream) or recompute them per reverse step: the
    same kernels on the same inputs, so the gradients are identical."""
    sc = terrain(44, 36, seed=25)
    tgt = _target(sc)
    kept = VF.adjoint_gradient(make_hip(sc), 14, tgt, checkpoint_every=4)
    runs = []
    for env in (("SFV_ADJOINT_OVERLAP", "0"), ("SFV_ADJOINT_KEEP", "0")):
        monkeypatch.setenv(*env)
        runs.append(VF.adjoint_gradient(make_hip(sc), 14, tgt, checkpoint_every=4))
    for again in runs:
        assert kept["J"] == again["J"] and kept["dJ_dCz"] == again["dJ_dCz"]
        assert np.array_equal(kept["dJ_dB"], again["dJ_dB"]) and np.array_equal(kept["dJ_dU0"], again["dJ_dU0"])


def test_adjoint_rejects_unsupported_runs():
    sc = terrain(20, 20, seed=1, bc=VF.NeumannBC())
    with pytest.raises(VF.ArgumentError):
        VF.adjoint_gradient(make_hip(sc), 3, np.zeros((20, 20)))


@pytest.mark.parametrize("nx,ny,gc,bc", [(3, 20, 2, "wall"), (26, 3, 2, "wall"), (4, 18, 3, "periodic"),
                                         (22, 17, 3, "wall"), (19, 23, 3, "periodic")])
def test_adjoint_narrow_grids_and_wide_halos(nx, ny, gc, bc):
    """gc = 3, and grids so narrow that one interior cell is the source of a
    low-side and a high-side ghost: the tails read the boundary condition's
    fold as a gather, which must add every ghost's adjoint into its source."""
    b = VF.WallBC() if bc == "wall" else VF.PeriodicBC()
    _check(smooth_flow(terrain(nx, ny, seed=31 + nx + ny, gc=gc, bc=b)), 8, 3)
