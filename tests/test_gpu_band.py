"""The march kernels' two face forms.  A wave evaluates its faces with the
ranged forms (shared reciprocals, the short square root) while every stage
input and bottom corner it reads is 0 or within [2^-100, 2^40] in magnitude;
when one is not, it runs its segment again with the general forms.  These
cases put such inputs in the middle of segments (tiny and subnormal momenta,
a tiny bottom corner), in one member of a batch, in Float32 storage, and turn
the ranged forms off altogether, and check every step bitwise against the
oracle."""
import numpy as np
import pytest

import volumefluxes as VF
from volumefluxes.scenarios import cell_bottom, terrain_initial_state, terrain_topography

from helpers import Scenario, _system, make_hip, make_oracle, terrain

pytestmark = pytest.mark.gpu


def _run(sc, nsteps, rows=12):
    """Step the march kernels (segment height forced) and the oracle; every
    dt and the final state must be equal."""
    print(f"[band] {sc.name} rows={rows} steps={nsteps}", flush=True)
    sim = make_hip(sc, "march")
    sim.set_march_rows(rows)
    o = make_oracle(sc)
    t = 0.0
    for k in range(nsteps):
        dr = o.step(float("inf"), t)
        dh = VF.perform_step(sim, float("inf"))
        assert dh == dr, f"dt differs at step {k}: {dh!r} vs {dr!r}"
        sim.t = float(sc.realtype(sc.realtype(sim.t) + sc.realtype(dh)))
        t = float(sc.realtype(sc.realtype(t) + sc.realtype(dr)))
    U_hip, U_ref = sim.get_full_state(), o.state()
    diff = np.argwhere(U_hip != U_ref)
    assert diff.size == 0, f"{len(diff)} cells differ, first {diff[:5].tolist()}"
    return U_ref


def _wet_cells(sc, count, seed):
    """Interior cells under at least 0.5 m of water, away from the edges."""
    g = sc.grid.gc
    depth = sc.U0[:, :, 0] - cell_bottom(sc.B)[g:-g, g:-g]
    nx, ny = depth.shape
    cand = np.argwhere(depth > 0.5)
    cand = cand[(cand[:, 0] > 8) & (cand[:, 0] < nx - 8) & (cand[:, 1] > 8) & (cand[:, 1] < ny - 8)]
    rng = np.random.default_rng(seed)
    return [tuple(c) for c in cand[rng.choice(len(cand), count, replace=False)]]


@pytest.mark.parametrize("value", [1e-40, -3e-35, 5e-324, -2.5e-310])
def test_tiny_momenta_mid_segment(value):
    """Momenta between 0 and 2^-100 (normal and subnormal) make the waves
    that read them fall back during their segments."""
    sc = terrain(150, 90, seed=41)
    U0 = sc.U0.copy()
    for k, (i, j) in enumerate(_wet_cells(sc, 6, seed=1)):
        U0[i, j, 1 + (k % 2)] = value
    sc.U0 = U0
    _run(sc, 12)


def test_tiny_bottom_corner():
    """Bottom corners of 1e-45 and -7e-38 (out of band) read mid-segment."""
    nx, ny, g = 128, 96, 2
    grid = VF.CartesianGrid(nx, ny, gc=g, boundary=VF.WallBC(), extent=[[0.0, nx * 1.0], [0.0, ny * 1.0]])
    B = terrain_topography(grid, seed=42, n_bumps=32, amp=(0.5, 5.0), width=(3.2, 16.0))
    B[g + 60, g + 40] = 1e-45
    B[g + 61, g + 41] = -7e-38
    system = _system(grid, B, sources=[VF.SourceTermBottom(), VF.ConstantRain(0.036)],
                     friction=VF.ImplicitFriction(Cz=0.03 * 0.03), kappa=1e-4)
    sc = Scenario("tiny_bottom", grid, system, VF.RungeKutta2(), 0.25, terrain_initial_state(grid, B, 2.0),
                  np.float64, B)
    _run(sc, 10)


def test_float32_storage_tiny_momenta():
    sc = terrain(96, 80, seed=43, realtype=np.float32)
    U0 = sc.U0.copy()
    for k, (i, j) in enumerate(_wet_cells(sc, 4, seed=2)):
        U0[i, j, 1 + (k % 2)] = np.float32(1e-40)      # a float32 subnormal
    sc.U0 = U0
    _run(sc, 10)


@pytest.mark.parametrize("rows", [1, 5, 12])
def test_ranged_forms_off(monkeypatch, rows):
    """SFV_RANGED=0: every wave takes the general forms from the start."""
    monkeypatch.setenv("SFV_RANGED", "0")
    _run(terrain(100, 70, seed=44), 10, rows=rows)


def test_batch_member_falls_back():
    """In a batch, one member's waves fall back and the others' do not; each
    member still ends where the oracle does."""
    scs = [terrain(90, 60, seed=45), terrain(120, 50, seed=46), terrain(64, 64, seed=47, bc=VF.PeriodicBC())]
    U0 = scs[1].U0.copy()
    for k, (i, j) in enumerate(_wet_cells(scs[1], 4, seed=3)):
        U0[i, j, 1 + (k % 2)] = 4e-39
    scs[1].U0 = U0
    sims = [make_hip(sc) for sc in scs]
    ens = VF.Ensemble(sims)
    ens.advance_steps(10)
    for sc, sim in zip(scs, sims):
        o = make_oracle(sc)
        t = 0.0
        for _ in range(10):
            t += o.step(float("inf"), t)
        assert sim.t == t, sc.name
        assert np.array_equal(sim.get_full_state(), o.state()), sc.name
    ens.close()
