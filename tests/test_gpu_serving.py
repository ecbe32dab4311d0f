"""Serving on the GPU: requests submitted while others run are admitted
under the HBM byte budget, and every result is bitwise the result of
running the request alone (simulate_to_time with the same arguments)."""
import time

import numpy as np
import pytest

import volumefluxes as VF

from helpers import dam_break, make_hip, terrain

pytestmark = pytest.mark.gpu


def _interior(sim):
    g = sim.grid.gc
    return sim.get_full_state()[g:-g, g:-g, :]


def _alone(sc, t_end, maximum_timestep=None):
    sim = make_hip(sc)
    VF.simulate_to_time(sim, t_end, maximum_timestep=maximum_timestep)
    return sim


def _requests():
    return [
        (terrain(40, 30, seed=1), 2.0),
        (terrain(130, 64, seed=2, bc=VF.PeriodicBC()), 1.0),
        (dam_break(96, 10, hr=0.0), 3.5),
        (terrain(64, 200, seed=3, friction=False), 0.75),
        (terrain(257, 129, seed=4), 1.25),
    ]


def test_served_results_equal_solo_runs():
    reqs = _requests()
    with VF.Server(8 << 30) as sv:
        ids = [sv.submit(sc.system, sc.grid, sc.U0, te, cfl=sc.cfl) for sc, te in reqs]
        for (sc, te), rid in zip(reqs, ids):
            info = sv.wait(rid, timeout=120)
            assert info["status"] == 2, info
            U, info = sv.fetch(rid)
            alone = _alone(sc, te)
            print(sc.name, info["t"], info["steps"], alone.t)
            assert info["t"] == alone.t == te
            assert np.array_equal(U, _interior(alone)), sc.name
        st = sv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0 and st["bytes_in_use"] == 0


def test_budget_limits_what_runs_at_once():
    reqs = _requests()
    per = [VF.build_config(sc.system, sc.stepper, sc.grid, sc.cfl, 0) for sc, _ in reqs]
    from volumefluxes import _capi as C
    import ctypes
    sizes = [C.lib().sfv_device_bytes_for(ctypes.byref(c)) for c in per]
    # at most about two at a time; the server holds each buffer rounded up to
    # its pool size class (at most 1/8 more, plus 4 KiB)
    budget = int(1.25 * max(sizes)) + min(sizes)
    with VF.Server(budget) as sv:
        ids = []
        for sc, te in reqs:
            ids.append(sv.submit(sc.system, sc.grid, sc.U0, te, cfl=sc.cfl))
            time.sleep(0.01)                  # arrivals while others run
        outs = [sv.fetch(rid) if sv.wait(rid, timeout=120) else None for rid in ids]
        st = sv.stats()
    assert st["bytes_peak"] <= budget
    for (sc, te), (U, info) in zip(reqs, outs):
        alone = _alone(sc, te)
        assert np.array_equal(U, _interior(alone)), sc.name
        assert info["admit_s"] >= info["submit_s"] and info["done_s"] >= info["admit_s"]


def test_request_over_budget_fails_with_budget_error():
    sc = terrain(256, 256, seed=9)
    with VF.Server(1 << 20) as sv:
        rid = sv.submit(sc.system, sc.grid, sc.U0, 1.0)
        info = sv.wait(rid, timeout=60)
        assert info["status"] == 3
        with pytest.raises(VF.BudgetError):
            sv.fetch(rid)


def test_unfusable_request_is_rejected_at_submit():
    sc = terrain(40, 40, seed=1, bc=VF.NeumannBC())
    with VF.Server(1 << 30) as sv:
        with pytest.raises(VF.ArgumentError):
            sv.submit(sc.system, sc.grid, sc.U0, 1.0)


def test_maximum_timestep_and_late_arrivals():
    a, b = terrain(90, 45, seed=11), terrain(45, 90, seed=12)
    with VF.Server(4 << 30) as sv:
        ra = sv.submit(a.system, a.grid, a.U0, 3.0, maximum_timestep=0.02)
        time.sleep(0.05)
        rb = sv.submit(b.system, b.grid, b.U0, 0.5)
        Ub, ib = sv.fetch(rb) if sv.wait(rb, timeout=60) else (None, None)
        Ua, ia = sv.fetch(ra) if sv.wait(ra, timeout=60) else (None, None)
    sa, sb = _alone(a, 3.0, maximum_timestep=0.02), _alone(b, 0.5)
    assert np.array_equal(Ua, _interior(sa)) and ia["t"] == 3.0
    assert np.array_equal(Ub, _interior(sb)) and ib["t"] == 0.5


def test_recycled_device_buffers_give_the_same_result():
    """A finished request's buffers go back to the server's pool and the next
    request of the same size takes them: its result must not see the old data."""
    a = terrain(120, 70, seed=31)
    b = dam_break(120, 70, hr=0.0)
    with VF.Server(2 << 30) as sv:
        outs = []
        for sc, te in [(a, 1.5), (b, 1.0), (a, 0.5)]:
            rid = sv.submit(sc.system, sc.grid, sc.U0, te, cfl=sc.cfl)
            assert sv.wait(rid, timeout=60)["status"] == 2
            outs.append(sv.fetch(rid)[0])
        st = sv.stats()
    assert st["host_admit_s"] >= 0 and st["host_retire_s"] >= 0 and st["busy_s"] > 0
    for (sc, te), U in zip([(a, 1.5), (b, 1.0), (a, 0.5)], outs):
        assert np.array_equal(U, _interior(_alone(sc, te))), sc.name


def test_large_and_small_grids_served_together():
    """Grids above 2^18 cells step in the large-grid engine, smaller ones in
    the small-grid engine (replayed several times per large replay); both
    give the results of the requests run alone."""
    reqs = [(terrain(600, 500, seed=51), 0.6), (terrain(48, 40, seed=52), 3.0),
            (dam_break(150, 20, hr=0.0), 2.0), (terrain(520, 640, seed=53, bc=VF.PeriodicBC()), 0.3)]
    with VF.Server(8 << 30) as sv:
        ids = []
        for sc, te in reqs:
            ids.append(sv.submit(sc.system, sc.grid, sc.U0, te, cfl=sc.cfl))
            time.sleep(0.005)
        outs = [sv.fetch(rid) if sv.wait(rid, timeout=120) else None for rid in ids]
        st = sv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for (sc, te), (U, info) in zip(reqs, outs):
        alone = _alone(sc, te)
        assert info["t"] == alone.t == te, sc.name
        assert np.array_equal(U, _interior(alone)), sc.name


def test_nonfinite_request_is_delivered_with_a_warning():
    """A NaN in the initial state makes dt NaN, as in the reference, which
    steps on with it (simulator.jl:190-192): the request finishes, its state
    is delivered with a RuntimeWarning, and it equals the solo run's."""
    sc = terrain(40, 30, seed=61)
    from volumefluxes.scenarios import cell_bottom
    depth = sc.U0[:, :, 0] - cell_bottom(sc.B)[2:-2, 2:-2]
    i, j = np.argwhere(depth > 0.5)[0]           # a wet cell: its NaN momentum reaches dt
    U0 = sc.U0.copy()
    U0[i, j, 1] = np.nan
    good = terrain(50, 20, seed=62)
    with VF.Server(1 << 30) as sv:
        rb = sv.submit(sc.system, sc.grid, U0, 2.0, cfl=sc.cfl)
        rg = sv.submit(good.system, good.grid, good.U0, 1.0, cfl=good.cfl)
        assert sv.wait(rb, timeout=60)["status"] == 2
        with pytest.warns(RuntimeWarning):
            Ub, ib = sv.fetch(rb)
        Ug, ig = sv.fetch(rg) if sv.wait(rg, timeout=60) else (None, None)
    solo = make_hip(sc)
    VF.set_current_state(solo, U0)
    with pytest.warns(RuntimeWarning):
        VF.simulate_to_time(solo, 2.0)
    a, b = Ub, _interior(solo)
    assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    assert np.array_equal(Ug, _interior(_alone(good, 1.0)))     # the other request is unaffected


def test_destroying_a_busy_server_fails_its_requests_cleanly():
    sc = terrain(300, 300, seed=63)
    sv = VF.Server(1 << 30)
    ids = [sv.submit(sc.system, sc.grid, sc.U0, 50.0, cfl=sc.cfl) for _ in range(3)]
    time.sleep(0.05)
    sv.close()                         # stops the worker: queued and running requests fail
    with VF.Server(1 << 30) as sv2:    # and the device is usable afterwards
        rid = sv2.submit(sc.system, sc.grid, sc.U0, 0.2, cfl=sc.cfl)
        assert sv2.wait(rid, timeout=60)["status"] == 2
        U, _ = sv2.fetch(rid)
    assert np.array_equal(U, _interior(_alone(sc, 0.2)))
    assert len(ids) == 3


@pytest.mark.timeout(600)
def test_serving_workload_mix_against_the_oracle():
    """The serving workload's request mix on a short horizon: Poisson arrivals
    at 50 per second (seeded), square grids from 64 to 1100 cells a side at
    dx = 1 m (half dam breaks 2 m | 1 m, half seeded terrains with rain and
    friction), t_end 10-15 s (the full workload draws U[10, 600] s).  The
    HBM budget holds the 1024^2 terrain or the 1100^2 dam break but not both,
    so the second waits in the queue until the first retires, and the requests
    behind it wait too (admission is in arrival order).  Three results (the
    1024^2 terrain, an 80^2 dam break and a 90^2 terrain) are compared bitwise
    with the oracle's simulate_to_time (simulator.jl:179-212) and the others
    with the same request run alone on the HIP backend."""
    import ctypes
    import threading

    import oracle as O
    from volumefluxes import _capi as C
    from volumefluxes.scenarios import serving_request

    spec = [(1024, "terrain", 11, 10.0, True), (1100, "dam_break", 0, 10.0, False),
            (80, "dam_break", 0, 10.0, True), (90, "terrain", 12, 10.0, True),
            (200, "terrain", 13, 12.0, False), (300, "dam_break", 0, 15.0, False),
            (64, "terrain", 5, 10.0, False), (512, "dam_break", 0, 11.0, False)]
    reqs = [serving_request(side, kind, seed) + (te, ora) for side, kind, seed, te, ora in spec]
    rng = np.random.default_rng(50)
    arrivals = np.cumsum(rng.exponential(1.0 / 50.0, len(reqs)))
    sizes = [C.lib().sfv_device_bytes_for(ctypes.byref(VF.build_config(s, VF.RungeKutta2(), g, 0.25, 0)))
             for g, s, _, _, _ in reqs]
    budget = int(1.2 * max(sizes[0], sizes[1])) + sum(sizes[2:])
    assert sizes[0] + sizes[1] > budget            # the two large grids cannot run together

    ids = [None] * len(reqs)
    with VF.Server(budget) as sv:
        def submitter():
            t_start = time.perf_counter()
            for k, (g, s, U0, te, _) in enumerate(reqs):
                delay = arrivals[k] - (time.perf_counter() - t_start)
                if delay > 0:
                    time.sleep(delay)
                ids[k] = sv.submit(s, g, U0, te)
        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        outs = []
        for k, rid in enumerate(ids):
            info = sv.wait(rid, timeout=300)
            assert info["status"] == 2, (k, info)
            outs.append(sv.fetch(rid))
        st = sv.stats()
    print(f"[serving mix] budget {budget / 1e6:.1f} MB, peak {st['bytes_peak'] / 1e6:.1f} MB, "
          f"replays {st['replays']}", flush=True)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["bytes_peak"] <= budget
    info0 = outs[0][1]
    for k in range(1, len(reqs)):
        assert outs[k][1]["admit_s"] >= info0["done_s"], k        # queued behind the budget
    for k, ((g, s, U0, te, ora), (U, info)) in enumerate(zip(reqs, outs)):
        assert info["t"] == te, k
        if ora:
            cfg = VF.build_config(s, VF.RungeKutta2(), g, 0.25, 0)
            o = O.Oracle2D(cfg, np.asfortranarray(s.equation.B.corners(g), dtype=np.float64))
            o.set_interior(U0)
            n = o.simulate_to_time(te)
            print(f"[serving mix] request {k} ({g.nx}^2): {info['steps']} steps served, {n} on the oracle",
                  flush=True)
            assert info["steps"] == n and o.t == info["t"], k
            assert np.array_equal(U, o.state()[2:-2, 2:-2, :]), k
        else:
            sim = VF.Simulator(VF.HIPBackend(np.float64), s, VF.RungeKutta2(), g, cfl=0.25)
            VF.set_current_state(sim, U0)
            VF.simulate_to_time(sim, te)
            assert sim.t == info["t"], k
            assert np.array_equal(U, _interior(sim)), k
