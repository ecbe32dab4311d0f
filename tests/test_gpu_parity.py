"""HIP backend vs the CPU oracle on identical inputs: FP64 states must be
bitwise equal (the per-step gate max|Δ|/max|ref| <= 1e-12 is then met with
zero), step by step, for the fused and per-pass kernels, on every boundary
condition, stepper, reconstruction and source combination the backend runs."""
import numpy as np
import pytest

import volumefluxes as VF

from helpers import dam_break, make_hip, make_oracle, max_rel_diff, run_both_compare, terrain

pytestmark = pytest.mark.gpu


KERNELS = ["march", "tiled", "hybrid", "per-pass"]


def _assert_bitwise(sc, nsteps, kernels=None, **kw):
    print(f"[parity] {sc.name} kernels={kernels} steps={nsteps}", flush=True)
    dts, U_hip, U_ref = run_both_compare(sc, nsteps, kernels=kernels, **kw)
    print(f"[parity]   dt0={dts[0]} rel={max_rel_diff(U_hip, U_ref):.3e}", flush=True)
    for k, (dh, dr) in enumerate(dts):
        assert dh == dr or (np.isnan(dh) and np.isnan(dr)), f"dt differs at step {k}: {dh!r} vs {dr!r}"
    diff = np.argwhere(U_hip != U_ref)
    assert diff.size == 0, f"{len(diff)} cells differ, first {diff[:5].tolist()}, rel={max_rel_diff(U_hip, U_ref)}"


@pytest.mark.parametrize("kernels", KERNELS)
def test_terrain_rk2_wall(kernels):
    _assert_bitwise(terrain(70, 45, seed=3), 30, kernels=kernels)


@pytest.mark.parametrize("kernels", KERNELS)
def test_terrain_rk2_periodic(kernels):
    _assert_bitwise(terrain(64, 16, seed=5, bc=VF.PeriodicBC()), 25, kernels=kernels)


@pytest.mark.parametrize("kernels", [None, "hybrid", "per-pass"])
def test_terrain_neumann_rk2(kernels):
    _assert_bitwise(terrain(33, 29, seed=7, bc=VF.NeumannBC()), 20, kernels=kernels)


@pytest.mark.parametrize("kernels", [None, "hybrid", "per-pass"])
def test_terrain_forward_euler(kernels):
    _assert_bitwise(terrain(40, 37, seed=9, stepper=VF.ForwardEulerStepper()), 30, kernels=kernels)


def test_default_kernels_by_configuration():
    """RK2 with interior ghost sources takes the march kernels; ForwardEuler
    and NeumannBC take the hybrid path (march stage-1 right-hand sides,
    per-pass updates)."""
    assert make_hip(terrain(20, 20, seed=1)).kernels == "march"
    assert make_hip(terrain(20, 20, seed=1, bc=VF.NeumannBC())).kernels == "hybrid"
    assert make_hip(terrain(20, 20, seed=1, stepper=VF.ForwardEulerStepper())).kernels == "hybrid"


@pytest.mark.parametrize("case", ["fe_periodic", "fe_neumann", "fe_norec", "neumann_f32", "fe_gc3", "fe_dry_front",
                                  "neumann_ragged"])
def test_hybrid_path_cases(case):
    FE = VF.ForwardEulerStepper()
    sc = {
        "fe_periodic": lambda: terrain(61, 45, seed=31, bc=VF.PeriodicBC(), stepper=FE),
        "fe_neumann": lambda: terrain(70, 33, seed=32, bc=VF.NeumannBC(), stepper=FE),
        "fe_norec": lambda: terrain(50, 20, seed=33, rec=VF.NoReconstruction(), stepper=FE),
        "neumann_f32": lambda: terrain(48, 40, seed=34, bc=VF.NeumannBC(), realtype=np.float32),
        "fe_gc3": lambda: terrain(40, 24, seed=35, gc=3, stepper=FE),
        "fe_dry_front": lambda: dam_break(128, 8, hl=0.005, hr=0.0, stepper=FE),
        "neumann_ragged": lambda: terrain(125, 37, seed=36, bc=VF.NeumannBC()),
    }[case]()
    _assert_bitwise(sc, 20, kernels="hybrid")


@pytest.mark.parametrize("shape", [(3, 2), (5, 3), (3, 70), (130, 3), (61, 17)])
def test_hybrid_small_and_ragged(shape):
    _assert_bitwise(terrain(*shape, seed=13, dx=1.0, n_bumps=3, stepper=VF.ForwardEulerStepper()), 12,
                    kernels="hybrid")


def test_no_reconstruction():
    _assert_bitwise(terrain(50, 20, seed=11, rec=VF.NoReconstruction()), 20)


@pytest.mark.parametrize("shape", [(1, 1), (3, 2), (5, 3), (3, 70), (130, 3), (65, 9), (61, 17), (121, 33)])
def test_small_and_ragged(shape):
    _assert_bitwise(terrain(*shape, seed=13, dx=1.0, n_bumps=3), 12)


@pytest.mark.parametrize("kernels", KERNELS)
def test_ragged_tiles(kernels):
    # widths and heights that leave partial tiles and partial wave strips
    _assert_bitwise(terrain(125, 37, seed=19), 10, kernels=kernels)


def test_gc3():
    _assert_bitwise(terrain(40, 24, seed=15, gc=3), 15)


@pytest.mark.parametrize("kernels", KERNELS)
def test_no_reconstruction_fused(kernels):
    _assert_bitwise(terrain(70, 30, seed=23, rec=VF.NoReconstruction()), 10, kernels=kernels)


def test_dam_break_dry_front():
    _assert_bitwise(dam_break(128, 8, hl=0.005, hr=0.0), 40)


def test_dam_break_y_direction():
    _assert_bitwise(dam_break(6, 150, hl=0.005, hr=0.001, direction=1), 40)


@pytest.mark.parametrize("kernels", KERNELS)
def test_terrain_f32_storage(kernels):
    _assert_bitwise(terrain(48, 40, seed=17, realtype=np.float32), 20, kernels=kernels)


@pytest.mark.parametrize("kernels", KERNELS)
def test_terrain_256_20_steps(kernels):
    _assert_bitwise(terrain(256, 256, seed=21), 20, kernels=kernels)


@pytest.mark.parametrize("rows", [1, 2, 5, 12, 40, 1024])
def test_segment_height_keeps_results_bitwise(rows):
    """The march kernels' segment height (autotuned or forced) changes which
    wave evaluates what, never a result."""
    sc = terrain(125, 37, seed=23)
    print(f"[parity] rows={rows}", flush=True)
    o = make_oracle(sc)
    sim = make_hip(sc, "march")
    sim.set_march_rows(rows)
    t = 0.0
    for k in range(8):
        dr = o.step(float("inf"), t)
        dh = VF.perform_step(sim, float("inf"))
        assert dh == dr, (k, dh, dr)
        t += dr
        sim.t = t
        assert np.array_equal(sim.get_full_state(), o.state()), (rows, k)
    assert sim.march_rows == rows


def test_model_height_at_1024_bitwise(monkeypatch):
    """1024^2 with the tuner off: the height model's choice (7 rows, 18 x 147
    waves, the last workgroup partly idle) against the oracle."""
    monkeypatch.setenv("SFV_AUTOTUNE", "0")
    sc = terrain(1024, 1024, seed=25)
    print("[parity] 1024x1024 model height", flush=True)
    o = make_oracle(sc)
    sim = make_hip(sc, "march")
    t = 0.0
    for k in range(3):
        dr = o.step(float("inf"), t)
        dh = VF.perform_step(sim, float("inf"))
        assert dh == dr, (k, dh, dr)
        t += dr
        sim.t = t
    assert sim.march_rows == 7
    assert np.array_equal(sim.get_full_state(), o.state())


def test_launch_of_several_rounds_bitwise():
    """600^2 at one-row segments: 6,000 waves, more workgroups than the device
    holds at once, numbered across the launch with idle waves in the last."""
    sc = terrain(600, 600, seed=24)
    print("[parity] 600x600 rows=1", flush=True)
    o = make_oracle(sc)
    sim = make_hip(sc, "march")
    sim.set_march_rows(1)
    t = 0.0
    for k in range(3):
        dr = o.step(float("inf"), t)
        dh = VF.perform_step(sim, float("inf"))
        assert dh == dr, (k, dh, dr)
        t += dr
        sim.t = t
    assert np.array_equal(sim.get_full_state(), o.state())


def test_autotuned_height_is_chosen_once_per_size_class():
    a = make_hip(terrain(90, 60, seed=29), "march")
    assert a.march_rows == 0
    VF.perform_step(a, float("inf"))
    r = a.march_rows
    assert 1 <= r <= 60
    b = make_hip(terrain(92, 61, seed=30), "march")       # same quarter-octave class
    VF.perform_step(b, float("inf"))
    assert b.march_rows == r
    sc = terrain(90, 60, seed=29)
    dts, U_hip, U_ref = run_both_compare(sc, 6, kernels="march")
    assert all(dh == dr for dh, dr in dts) and np.array_equal(U_hip, U_ref)


def test_hybrid_path_height_tuned_and_results_height_independent():
    """ForwardEuler / NeumannBC handles tune the march stage-1 height on their
    first step like march handles, and every height gives the oracle's bits."""
    sc = terrain(90, 60, seed=37, stepper=VF.ForwardEulerStepper())
    a = make_hip(sc)
    VF.perform_step(a, float("inf"))
    assert a.kernels == "hybrid" and a.march_rows >= 1
    for rows in (1, 7, 64):
        sim = make_hip(sc)
        sim.set_march_rows(rows)
        o = make_oracle(sc)
        t = 0.0
        for k in range(10):
            dr = o.step(float("inf"), t)
            dh = VF.perform_step(sim, float("inf"))
            assert dh == dr, (rows, k)
            t += dr
            sim.t = t
        assert np.array_equal(sim.get_full_state(), o.state()), rows


def test_tuner_may_pick_tiled_kernels_and_results_stay_bitwise():
    """A handle whose variant was not set times the LDS-tiled kernels next to
    the march heights on its first step and takes the faster; a handle set to
    "march" keeps it.  Either way the results are the oracle's bits."""
    sc = terrain(300, 280, seed=41)
    dts, U_hip, U_ref = run_both_compare(sc, 8)
    assert all(dh == dr for dh, dr in dts) and np.array_equal(U_hip, U_ref)
    auto = make_hip(sc)
    VF.perform_step(auto, float("inf"))
    assert auto.kernels in ("march", "tiled")
    assert auto.march_rows >= 1
    forced = make_hip(sc, "march")
    VF.perform_step(forced, float("inf"))
    assert forced.kernels == "march"
    assert np.array_equal(auto.get_full_state(), forced.get_full_state())


def test_switching_kernel_variants_mid_run_keeps_oracle_bits():
    """The march and tiled paths read ghost cells through the boundary maps
    and fill them only when a host call ends; the per-pass and hybrid paths
    read them from memory.  Switching between them between calls must give
    the oracle's bits step for step."""
    import ctypes
    from volumefluxes import _capi as C
    sc = terrain(130, 70, seed=61)
    sim = make_hip(sc)
    o = make_oracle(sc)
    t = 0.0
    for variant in ("march", "per-pass", "tiled", "hybrid", "march", "tiled", "per-pass"):
        sim.set_kernels(variant)
        for _ in range(4):
            dr = o.step(float("inf"), t)
            dh = VF.perform_step(sim, float("inf"))
            assert dh == dr, variant
            t += dr
            sim.t = t
        assert np.array_equal(sim.get_full_state(), o.state()), variant
    # a device-resident run, then a timing call, then the per-pass path again
    sim.set_kernels("march")
    VF.advance_steps(sim, 6)
    for _ in range(6):
        t += o.step(float("inf"), t)
    assert np.array_equal(sim.get_full_state(), o.state())
    before = sim.get_full_state()
    out = ctypes.c_double()
    C.check(sim.handle, C.lib().sfv_time_kernel(sim.handle, b"step", 3, ctypes.byref(out)))
    assert out.value > 0.0
    assert np.array_equal(sim.get_full_state(), before)       # timing leaves the state alone
    sim.set_kernels("per-pass")
    for _ in range(3):
        dr = o.step(float("inf"), t)
        dh = VF.perform_step(sim, float("inf"))
        assert dh == dr
        t += dr
    assert np.array_equal(sim.get_full_state(), o.state())


def _random_case(k):
    rng = np.random.default_rng(9000 + k)
    gc = int(rng.choice([2, 3]))
    nx, ny = int(rng.integers(max(gc, 3), 140)), int(rng.integers(max(gc, 3), 90))
    bc = [VF.WallBC(), VF.PeriodicBC(), VF.NeumannBC()][int(rng.integers(0, 3))]
    stepper = VF.RungeKutta2() if rng.random() < 0.6 else VF.ForwardEulerStepper()
    rec = VF.NoReconstruction() if rng.random() < 0.2 else None
    sc = terrain(nx, ny, seed=9100 + k, gc=gc, bc=bc, stepper=stepper, rec=rec,
                 friction=bool(rng.random() < 0.7), rain=0.036 if rng.random() < 0.6 else 0.0,
                 n_bumps=int(rng.integers(1, 6)))
    kernels = None if rng.random() < 0.75 else "per-pass"
    return sc, kernels


@pytest.mark.parametrize("k", range(24))
def test_random_configurations_bitwise(k):
    """Seeded random sizes, ghost widths, boundary conditions, steppers,
    reconstructions and sources, on the default kernel path (march, tiled or
    hybrid as the handle picks) or the per-pass one: 8 steps, bitwise
    against the oracle, ghost cells included."""
    sc, kernels = _random_case(k)
    _assert_bitwise(sc, 8, kernels=kernels)
