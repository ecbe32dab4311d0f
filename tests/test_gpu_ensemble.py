"""Ensembles on the GPU: every member of a batch advanced together must end
bitwise where it ends when advanced alone (which the parity tests tie to the
oracle), with its own dt and time, over uneven grid sizes, boundary
conditions, sources and friction laws."""
import numpy as np
import pytest

import volumefluxes as VF

from helpers import dam_break, make_hip, make_oracle, terrain

pytestmark = pytest.mark.gpu


def _mix(realtype=np.float64):
    """Uneven members: tiny to multi-tile, wall and periodic, with and
    without friction and rain, a wet/dry dam break."""
    return [
        terrain(37, 23, seed=3, realtype=realtype),
        terrain(64, 64, seed=4, bc=VF.PeriodicBC(), realtype=realtype),
        terrain(130, 70, seed=5, friction=False, realtype=realtype),
        terrain(61, 300, seed=6, rain=0.0, realtype=realtype),
        terrain(257, 129, seed=7, realtype=realtype),
        dam_break(90, 12, hl=0.005, hr=0.0) if realtype == np.float64 else terrain(90, 12, seed=8, realtype=realtype),
    ]


def _solo(sc, nsteps):
    sim = make_hip(sc)
    VF.advance_steps(sim, nsteps)
    return sim


@pytest.mark.parametrize("realtype", [np.float64, np.float32])
def test_ensemble_steps_equal_solo_runs(realtype):
    scs = _mix(realtype)
    sims = [make_hip(sc) for sc in scs]
    ens = VF.Ensemble(sims)
    steps = ens.advance_steps(25)
    assert list(steps) == [25] * len(scs)
    for sc, sim in zip(scs, sims):
        solo = _solo(sc, 25)
        print(sc.name, sim.t, solo.t)
        assert sim.t == solo.t, sc.name
        assert sim.current_timestep == solo.current_timestep, sc.name
        assert np.array_equal(sim.get_full_state(), solo.get_full_state()), sc.name
    ens.close()


def test_ensemble_matches_oracle_small_members():
    scs = [terrain(37, 23, seed=11), terrain(50, 41, seed=12, bc=VF.PeriodicBC()), dam_break(64, 9, hr=0.0)]
    sims = [make_hip(sc) for sc in scs]
    ens = VF.Ensemble(sims)
    ens.advance_steps(12)
    for sc, sim in zip(scs, sims):
        o = make_oracle(sc)
        t = 0.0
        for _ in range(12):
            t += o.step(float("inf"), t)
        assert sim.t == t, sc.name
        assert np.array_equal(sim.get_full_state(), o.state()), sc.name


def test_ensemble_runs_continue_and_interleave_with_solo_steps():
    scs = [terrain(70, 33, seed=21), terrain(33, 70, seed=22)]
    sims = [make_hip(sc) for sc in scs]
    ens = VF.Ensemble(sims)
    ens.advance_steps(5)
    VF.advance_steps(sims[0], 3)          # a member stepped on its own between batch runs
    ens.advance_steps(7)
    solo0, solo1 = _solo(scs[0], 15), _solo(scs[1], 12)
    assert np.array_equal(sims[0].get_full_state(), solo0.get_full_state())
    assert np.array_equal(sims[1].get_full_state(), solo1.get_full_state())
    assert sims[0].t == solo0.t and sims[1].t == solo1.t


def test_ensemble_simulate_to_time_per_member_endtimes():
    scs = [terrain(48, 40, seed=31), terrain(96, 20, seed=32), dam_break(80, 10, hr=0.0001)]
    endtimes = [1.5, 4.0, 2.25]
    sims = [make_hip(sc) for sc in scs]
    ens = VF.Ensemble(sims)
    steps = ens.simulate_to_time(endtimes, maximum_timestep=0.05)
    for sc, sim, te, ns in zip(scs, sims, endtimes, steps):
        solo = make_hip(sc)
        VF.simulate_to_time(solo, te, maximum_timestep=0.05)
        assert sim.t == solo.t == te, sc.name
        assert np.array_equal(sim.get_full_state(), solo.get_full_state()), sc.name
        assert ns > 0


def test_ensemble_many_members_and_a_large_one():
    rng = np.random.default_rng(5)
    scs = [terrain(int(rng.integers(20, 90)), int(rng.integers(20, 90)), seed=100 + k) for k in range(40)]
    scs.append(terrain(1024, 700, seed=99))
    sims = [make_hip(sc) for sc in scs]
    ens = VF.Ensemble(sims)
    ens.advance_steps(6)
    for k in (0, 7, 19, 39, 40):
        solo = _solo(scs[k], 6)
        assert np.array_equal(sims[k].get_full_state(), solo.get_full_state()), scs[k].name


def test_ensemble_rejects_unfusable_members():
    sc = terrain(40, 40, seed=1, bc=VF.NeumannBC())
    sim = make_hip(sc)
    with pytest.raises(VF.ArgumentError):
        VF.Ensemble([sim])
    sc2 = terrain(40, 40, seed=1, stepper=VF.ForwardEulerStepper())
    with pytest.raises(VF.ArgumentError):
        VF.Ensemble([make_hip(sc2)])


def test_members_with_different_segment_heights():
    """Each member keeps its own march segment height inside the batch."""
    scs = [terrain(70, 45, seed=41), terrain(33, 90, seed=42), terrain(128, 20, seed=43, bc=VF.PeriodicBC())]
    sims = [make_hip(sc) for sc in scs]
    for sim, rows in zip(sims, (1, 3, 17)):
        sim.set_march_rows(rows)
    ens = VF.Ensemble(sims)
    ens.advance_steps(10)
    for sc, sim in zip(scs, sims):
        solo = _solo(sc, 10)
        assert sim.t == solo.t, sc.name
        assert np.array_equal(sim.get_full_state(), solo.get_full_state()), sc.name
    ens.close()


def test_batch_height_tuning_keeps_results(monkeypatch):
    """A batch run times the height rule's choice against its neighbours on
    its first replays and keeps the fastest.  Heights never change results:
    the tuned batch ends bitwise where an untuned batch and the solo runs
    end, and a later run with the same member sizes keeps the choice."""
    scs = [terrain(70, 45, seed=51), terrain(200, 90, seed=52), terrain(33, 130, seed=53),
           dam_break(90, 12, hl=0.005, hr=0.0)]
    sims = [make_hip(sc) for sc in scs]
    ens = VF.Ensemble(sims)
    assert ens.march_rows == 0
    ens.advance_steps(48)                 # 6 replays: at most 5 timed trials, then the kept height
    rows = ens.march_rows
    assert rows in (1, 2, 3, 4, 6, 8, 12, 16, 24)
    ens.advance_steps(16)
    assert ens.march_rows == rows
    monkeypatch.setenv("SFV_AUTOTUNE", "0")
    sims0 = [make_hip(sc) for sc in scs]
    ens0 = VF.Ensemble(sims0)
    ens0.advance_steps(64)
    for sc, sim, sim0 in zip(scs, sims, sims0):
        assert sim.t == sim0.t, sc.name
        assert np.array_equal(sim.get_full_state(), sim0.get_full_state()), sc.name
        solo = _solo(sc, 64)
        assert np.array_equal(sim.get_full_state(), solo.get_full_state()), sc.name
    ens.close()
    ens0.close()
