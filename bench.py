#!/usr/bin/env python3
"""Benchmark: 2D shallow-water RK2 step on one MI355X (gfx950).

Workload (the single-GPU terrain run): 2048 x 2048 cells at dx = dy = 1 m,
bottom = 1e-3 slope + 32 seeded Gaussian bumps (amplitude U[0.5, 5] m, width
U[20, 200] m), initial surface w = max(B_cell, 2 m) so cells with B > 2 m
start dry, Manning n = 0.03 (bh2021 law, Cz = n^2), uniform rain 1e-5 m/s
(ConstantRain(0.036) m/h), wall boundaries, desingularisation kappa = 1e-4,
depth cutoff 1e-5, RungeKutta2, CFL 0.25.  A "step" is one full RK2 step of
the whole grid; the timed region runs --steps steps device-resident with the
state already in HBM.

Prints one JSON line: throughput (cell-updates/s), the roofline of the
dominant kernel from HIP-event timings on the kernel's own stream, and the
CPU oracle timed on this host's cores on a bounded sample of the same run.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "volumefluxes.jl_amd"))

import ctypes  # noqa: E402

import numpy as np  # noqa: E402

import volumefluxes as VF  # noqa: E402
from volumefluxes import _capi as C  # noqa: E402
from volumefluxes.scenarios import (serving_requests, terrain_initial_state, terrain_topography,  # noqa: E402
                                   terrain_workload)

HBM_PEAK_GBS = 8000.0                 # MI355X HBM3E, 8 TB/s
BYTES_PER_CELL_STEP = {"f64": 136, "f32": 80}   # compulsory HBM bytes per RK2 cell-update (B stays f64)
BYTES_PER_CELL_STAGE = {"f64": {"stage1": 56, "stage2": 80}, "f32": {"stage1": 36, "stage2": 44}}
METRIC = "cell-updates/s + HBM GB/s % peak (2D SWE RK2 FP64); ensemble p50/p99 vs plain"


def build_scenario(n: int, precision: str, seed: int):
    return terrain_workload(n, precision, seed)


def read_traffic(kernel: str, workload: str):
    """Measured HBM bytes per launch of `kernel` from the committed rocprofv3
    PMC summary (FETCH_SIZE, WRITE_SIZE, calibrated against a copy kernel)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if parts and parts[0] == kernel:
            kv = dict(p.split("=") for p in parts[1:] if "=" in p)
            if "bytes_per_launch" in kv:
                return float(kv["bytes_per_launch"])
    return None


VALU_WAVE_INSTR_PER_S = 256 * 2.4e9   # MI355X: 256 CUs, each issuing one 64-lane VALU instruction per clock


def read_valu(kernel: str, workload: str):
    """SQ_INSTS_VALU (64-lane wave-instructions per launch) of `kernel` from the
    committed rocprofv3 PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", f"valu_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if parts and parts[0] == kernel:
            kv = dict(p.split("=") for p in parts[1:] if "=" in p)
            if "SQ_INSTS_VALU" in kv and "dispatches" in kv:
                return float(kv["SQ_INSTS_VALU"])
    return None


def host_cpu():
    """(logical CPUs this process may use, CPU model name) of this host."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None and os.path.exists("/proc/cpuinfo"):
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    return len(os.sched_getaffinity(0)), model


def cpu_stoker_1d(threads: int):
    """The 1D CPU workload on the oracle: Stoker dam break, 1024 cells on
    [0, 1000] m, flat bottom, h = 2 m | 1 m, wall BCs, RK2, CFL 0.25, to t = 20 s
    with match_endtime.  Median of 3 whole runs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the reference restatement: CPU baseline only
    O.set_threads(threads)
    g1 = VF.CartesianGrid(1024, gc=2, boundary=VF.WallBC(), extent=[[0.0, 1000.0]])
    eq1 = VF.ShallowWaterEquations1D(VF.ConstantBottomTopography())
    s1 = VF.ConservedSystem(None, VF.LinearReconstruction(), VF.CentralUpwind(eq1), eq1, g1, [], None)
    g2, s2 = VF.embed_1d(s1, g1)
    cfg = VF.build_config(s2, VF.RungeKutta2(), g2, 0.25, C.SFV_F64)
    cfg.ny, cfg.dy = 1, cfg.dx
    U1 = np.zeros((1024, 2))
    U1[:, 0] = np.where(g1.cell_centers() < 500.0, 2.0, 1.0)
    times, steps = [], 0
    for _ in range(3):
        o = O.Oracle1D(cfg, np.ascontiguousarray(s1.equation.B.corners(g1), dtype=np.float64))
        o.set_interior(U1)
        t0 = time.perf_counter()
        steps = o.simulate_to_time(20.0)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": 1024 * steps / med, "unit": "cell-updates/s", "steps": int(steps), "seconds": med,
            "threads": threads, "workload": "1D dam break, 1024 cells on [0, 1000] m, RK2, CFL 0.25, to t = 20 s"}


def cpu_baseline(grid, system, realtype, U0, B, steps_per_run: int, runs: int, threads: int):
    """The CPU oracle (a loop-faithful restatement of the reference CPU path)
    on this host: 1 warm-up step, then the median of `runs` timed runs of
    `steps_per_run` steps, OpenMP over grid rows on `threads` threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the reference restatement: CPU baseline only, never the measured product
    O.set_threads(threads)
    cfg = VF.build_config(system, VF.RungeKutta2(), grid, 0.25, C.SFV_F64 if realtype == np.float64 else 1)
    o = O.Oracle2D(cfg, B)
    o.set_interior(U0)
    t = 0.0
    t += o.step(float("inf"), t)
    first = o.state()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        for _ in range(steps_per_run):
            t += o.step(float("inf"), t)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return first, grid.nx * grid.ny * steps_per_run / med, med


def oracle_first_step(grid, system, realtype, U0, B):
    """The oracle's state after one step from U0 (the parity probe's reference)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the reference restatement, used only to check results
    O.set_threads(os.cpu_count() or 1)
    cfg = VF.build_config(system, VF.RungeKutta2(), grid, 0.25, C.SFV_F64 if realtype == np.float64 else 1)
    o = O.Oracle2D(cfg, B)
    o.set_interior(U0)
    o.step(float("inf"), 0.0)
    return o.state()


ENSEMBLE_SIDES = (128, 256, 512, 768, 1024)


def ensemble_scenarios(members: int, seed: int):
    """The ensemble workload: grid sides drawn per instance (and per axis)
    from ENSEMBLE_SIDES, each instance with its own seeded bumps, rain rate
    U[0, 5e-5] m/s and Manning n U[0.01, 0.05] (Cz = n^2); f64, RK2, walls."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(members):
        nx, ny = (int(v) for v in rng.choice(ENSEMBLE_SIDES, 2))
        rain = float(rng.uniform(0.0, 5e-5))
        n_manning = float(rng.uniform(0.01, 0.05))
        grid = VF.CartesianGrid(nx, ny, gc=2, boundary=VF.WallBC(), extent=[[0.0, float(nx)], [0.0, float(ny)]])
        B = terrain_topography(grid, seed=seed * 1000 + k, n_bumps=32, slope=1e-3, amp=(0.5, 5.0),
                               width=(20.0, 200.0))
        eq = VF.ShallowWaterEquations(VF.BottomTopography2D(B, None, grid), desingularizing_kappa=1e-4)
        system = VF.ConservedSystem(None, VF.LinearReconstruction(), VF.CentralUpwind(eq), eq, grid,
                                    [VF.SourceTermBottom(), VF.ConstantRain(rain * 3600)],
                                    VF.ImplicitFriction(Cz=n_manning * n_manning))
        out.append((grid, system, terrain_initial_state(grid, B, 2.0)))
    return out


def ensemble_bench(members: int, steps: int, seed: int):
    """Batched (one persistent launch per RK stage for all members, graph
    replays) against plain (the same instances one at a time, each a
    device-resident advance_steps), all submitted at t = 0.  Time-to-result
    of an instance is when its final state is ready; p50/p99 over instances."""
    scen = ensemble_scenarios(members, seed)
    backend = VF.make_hip_backend(np.float64)

    def sims():
        out = []
        for grid, system, U0 in scen:
            sim = VF.Simulator(backend, system, VF.RungeKutta2(), grid, cfl=0.25)
            VF.set_current_state(sim, U0)
            out.append(sim)
        return out

    L = C.lib()
    cells = sum(g.nx * g.ny for g, _, _ in scen)
    # warm-up of both paths on throw-away copies of the first members
    warm = sims()[:8]
    for sim in warm:
        VF.advance_steps(sim, 2)
    ew = VF.Ensemble(warm)
    ew.advance_steps(2)
    ew.close()
    del warm

    plain = sims()
    for sim in plain:
        C.check(sim.handle, L.sfv_synchronize(sim.handle))
    done_plain = []
    t0 = time.perf_counter()
    for sim in plain:
        VF.advance_steps(sim, steps)
        done_plain.append(time.perf_counter() - t0)
    plain_s = done_plain[-1]

    batched = sims()
    for sim in batched:
        C.check(sim.handle, L.sfv_synchronize(sim.handle))
    ens = VF.Ensemble(batched)
    t0 = time.perf_counter()
    ens.advance_steps(steps)
    batched_s = time.perf_counter() - t0
    done_batched = [batched_s] * members

    picks = sorted(set([0, members // 3, members // 2, members - 1] + list(range(0, members, max(1, members // 8)))))
    same = all(np.array_equal(plain[k].get_full_state(), batched[k].get_full_state()) and plain[k].t == batched[k].t
               for k in picks)
    batch_rows = ens.march_rows
    ens.close()

    def pct(v, q):
        return float(np.percentile(np.asarray(v), q))
    return {
        "workload": (f"{members} instances, nx and ny each drawn from {list(ENSEMBLE_SIDES)}, own seeded bumps, "
                     f"rain U[0,5e-5] m/s, Manning n U[0.01,0.05], f64 RK2, {steps} steps each, own dt"),
        "cells_total": cells,
        "batched": {"cell_updates_per_s": cells * steps / batched_s, "seconds": batched_s,
                    "p50_s": pct(done_batched, 50), "p99_s": pct(done_batched, 99),
                    "march_rows": batch_rows},
        "plain": {"cell_updates_per_s": cells * steps / plain_s, "seconds": plain_s,
                  "p50_s": pct(done_plain, 50), "p99_s": pct(done_plain, 99)},
        "throughput_ratio": plain_s / batched_s,
        "p99_ratio": pct(done_plain, 99) / pct(done_batched, 99),
        "parity_batched_vs_plain": ("bitwise" if same else "MISMATCH") + f" ({len(picks)} members compared)",
    }


def serving_scenarios(rate: float, duration: float, t_range, seed: int):
    return serving_requests(rate, duration, t_range, seed)


def _progress(msg: str) -> None:
    """A progress line on stderr (stdout carries the one JSON result line)."""
    print(msg, file=sys.stderr, flush=True)


def serving_bench(rate: float, duration: float, t_range, budget_gb: float, seed: int, cap_s: float = 0.0):
    """Continuous batching (Server: arrivals submitted in real time, admitted
    under the HBM budget, advanced together) against plain one-at-a-time
    FIFO service of the same requests (each a device-resident
    simulate_to_time; its queue is replayed from the measured service times
    and the same arrival times).  Time-to-result = completion - arrival; it
    splits into the queue wait (admission - arrival) and the time on the
    device (completion - admission).

    With cap_s > 0 (a workload beyond what the device can finish in a
    measurement) each path runs for cap_s seconds from the first arrival and
    stops: the figures are the cell-updates delivered in that time (finished
    requests and the progress of running ones), the requests finished, and
    time-to-result over the finished ones; the others are reported as
    unfinished."""
    import threading
    reqs = serving_scenarios(rate, duration, t_range, seed)
    n = len(reqs)
    budget = int(budget_gb * (1 << 30))
    backend = VF.make_hip_backend(np.float64)
    # warm-up of both paths
    w = reqs[0]
    with VF.Server(budget) as sv:
        rid = sv.submit(w["system"], w["grid"], w["U0"], 0.1)
        sv.wait(rid, timeout=120)
        sv.fetch(rid)
    sim = VF.Simulator(backend, w["system"], VF.RungeKutta2(), w["grid"])
    VF.set_current_state(sim, w["U0"])
    VF.simulate_to_time(sim, 0.1)
    del sim
    _progress(f"[serving] {n} requests over {duration} s of arrivals, t_end U{tuple(t_range)}, budget {budget_gb} GB")

    # plain: one request at a time in arrival order, each a device-resident
    # simulate_to_time (in chunks of steps, so a cap can stop a long one);
    # the FIFO clock starts at the first arrival and idles until the next
    # arrival when the queue is empty
    service, solo_states, cells = [], {}, []
    picks = set(range(0, n, max(1, n // 6))) | {n - 1}
    t_plain = time.perf_counter()
    clock, done, plain_cells = reqs[0]["arrival"], [], 0.0
    L0 = C.lib()
    for k, r in enumerate(reqs):
        if cap_s > 0 and clock >= cap_s:
            break
        clock = max(clock, r["arrival"])
        sim = VF.Simulator(backend, r["system"], VF.RungeKutta2(), r["grid"])
        VF.set_current_state(sim, r["U0"])
        C.check(sim.handle, L0.sfv_synchronize(sim.handle))
        tt = ctypes.c_double(0.0)
        steps, t0, finished = 0, time.perf_counter(), False
        while True:
            st_ = ctypes.c_int64()
            chunk = 0 if cap_s <= 0 else 20000
            C.check(sim.handle, L0.sfv_run(sim.handle, r["t_end"], 1, 0.0, chunk, ctypes.byref(tt), ctypes.byref(st_)))
            steps += st_.value
            if tt.value >= r["t_end"] or st_.value == 0:
                finished = True
                break
            if cap_s > 0 and clock + (time.perf_counter() - t0) >= cap_s:
                break
        el = time.perf_counter() - t0
        clock += el
        plain_cells += r["grid"].nx * r["grid"].ny * steps
        if finished:
            service.append(el)
            cells.append(r["grid"].nx * r["grid"].ny * steps)
            done.append(clock)
            if k in picks:
                g = r["grid"].gc
                solo_states[k] = sim.get_full_state()[g:-g, g:-g, :]
        del sim
        if (k + 1) % 25 == 0 or k + 1 == n:
            _progress(f"[serving] plain: {k + 1}/{n} requests run alone, {time.perf_counter() - t_plain:.0f} s")
    plain_done = len(done)
    plain_lat = [d - r["arrival"] for d, r in zip(done, reqs)]
    plain_span = (cap_s if cap_s > 0 else done[-1]) - reqs[0]["arrival"]

    # served: real-time arrivals into the server
    ids = [None] * n
    with VF.Server(budget) as sv:
        def submitter():
            t_start = time.perf_counter()
            for k, r in enumerate(reqs):
                delay = r["arrival"] - (time.perf_counter() - t_start)
                if delay > 0:
                    time.sleep(delay)
                ids[k] = sv.submit(r["system"], r["grid"], r["U0"], r["t_end"])
        t_served = time.perf_counter()
        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        infos, same, served_cells, unfinished, compared = [], True, 0.0, 0, 0
        for k, rid in enumerate(ids):
            while True:
                left = cap_s - (time.perf_counter() - t_served) if cap_s > 0 else 30.0
                if left <= 0:
                    break
                try:
                    sv.wait(rid, timeout=min(30.0, left))
                    break
                except TimeoutError:
                    st = sv.stats()
                    _progress(f"[serving] served: {st['completed']}/{n} done, {st['running']} running, "
                              f"{st['queued']} queued, {st['bytes_in_use'] / 1e9:.2f} GB in use, "
                              f"{time.perf_counter() - t_served:.0f} s")
            info = sv.poll(rid)
            served_cells += info["cell_updates"]
            if info["status"] != 2:
                unfinished += 1
                continue
            U, info = sv.fetch(rid)
            infos.append(info)
            if k in solo_states:
                same = same and np.array_equal(U, solo_states[k])
                compared += 1
        st = sv.stats()
        span_cap = time.perf_counter() - t_served
    _progress(f"[serving] served: {len(infos)} of {n} done in {span_cap:.0f} s")
    lat = [i["done_s"] - i["submit_s"] for i in infos]
    wait = [i["admit_s"] - i["submit_s"] for i in infos]
    on_dev = [i["done_s"] - i["admit_s"] for i in infos]
    span = (span_cap if cap_s > 0 else max(i["done_s"] for i in infos) - min(i["submit_s"] for i in infos))
    total = float(sum(cells)) if cap_s <= 0 else served_cells

    def pct(v, q):
        return float(np.percentile(np.asarray(v), q)) if len(v) else None
    served_rate = total / span
    plain_rate = (plain_cells if cap_s > 0 else total) / plain_span
    if cap_s > 0:
        return {
            "workload": (f"Poisson arrivals {rate}/s for {duration} s ({n} requests), square grids side log-uniform "
                         f"[64, 2048], half dam breaks / half terrains, f64 RK2, t_end U[{t_range[0]}, {t_range[1]}] s, "
                         f"HBM budget {budget_gb} GB; each path stopped {cap_s:.0f} s after the first arrival"),
            "requests": n,
            "load": (f"offered cell-updates per second of arrivals: {'' if unfinished == 0 else 'at least '}"
                     f"{served_cells / duration:.3g} (the cell-updates the requests{'' if unfinished == 0 else ' finished or in progress'}"
                     f" needed, over the {duration} s arrival window), {span / duration:.0f}x the served rate; in "
                     f"{cap_s:.0f} s the served path finished {len(infos)} and the plain FIFO path {plain_done} of {n}"),
            "served": {"finished": len(infos), "unfinished": unfinished,
                       "cell_updates_per_s": served_rate, "cell_updates": served_cells,
                       "p50_s_finished": pct(lat, 50), "p99_s_finished": pct(lat, 99),
                       "queue_wait_p50_s": pct(wait, 50), "queue_wait_p99_s": pct(wait, 99),
                       "on_device_p50_s": pct(on_dev, 50), "on_device_p99_s": pct(on_dev, 99),
                       "busy_s": st["busy_s"], "peak_bytes": st["bytes_peak"], "budget_bytes": budget,
                       "replays": st["replays"]},
            "plain_fifo": {"finished": plain_done, "cell_updates_per_s": plain_rate, "cell_updates": plain_cells,
                           "p50_s_finished": pct(plain_lat, 50), "p99_s_finished": pct(plain_lat, 99)},
            "throughput_ratio": served_rate / plain_rate,
            "parity_served_vs_alone": ("bitwise" if same else "MISMATCH") + f" ({compared} requests compared)",
        }
    return {
        "workload": (f"Poisson arrivals {rate}/s for {duration} s ({n} requests), square grids side log-uniform "
                     f"[64, 2048], half dam breaks / half terrains, f64 RK2, t_end U[{t_range[0]}, {t_range[1]}] s, "
                     f"HBM budget {budget_gb} GB"),
        "requests": n,
        "cell_updates_total": total,
        "offered_cell_updates_per_s": total / duration,
        "offered_over_served_capacity": (total / duration) / served_rate,
        "served": {"p50_s": pct(lat, 50), "p99_s": pct(lat, 99), "queue_wait_p50_s": pct(wait, 50),
                   "queue_wait_p99_s": pct(wait, 99), "on_device_p50_s": pct(on_dev, 50),
                   "on_device_p99_s": pct(on_dev, 99), "cell_updates_per_s": served_rate, "makespan_s": span,
                   "busy_s": st["busy_s"], "host_admit_s": st["host_admit_s"], "host_retire_s": st["host_retire_s"],
                   "peak_bytes": st["bytes_peak"], "budget_bytes": budget, "replays": st["replays"]},
        "plain_fifo": {"p50_s": pct(plain_lat, 50), "p99_s": pct(plain_lat, 99),
                       "cell_updates_per_s": total / plain_span, "makespan_s": plain_span,
                       "service_s_total": float(sum(service))},
        "throughput_ratio": served_rate / plain_rate,
        "p99_ratio": pct(plain_lat, 99) / pct(lat, 99),
        "parity_served_vs_alone": ("bitwise" if same else "MISMATCH") + f" ({compared} requests compared)",
    }


def adjoint_bench(seed: int, n: int = 512, steps: int = 200, checkpoint_every: int = 20):
    """Gradient of J = sum (h_T - h_target)^2 with respect to every bottom
    corner and Manning's n on the n x n terrain (f64, RK2, `steps` steps),
    h_target from a forward run on a perturbed topography.  Checked by
    central finite differences on seeded corners and on n, and against the
    oracle's forward-mode derivative along a random direction."""
    from volumefluxes.scenarios import cell_bottom
    grid, system, _, _, B = build_scenario(n, "f64", seed)
    backend = VF.make_hip_backend(np.float64)
    g = grid.gc
    # The bumps of the single-grid workload crowd a smaller domain: at 2 m the
    # 512 m terrain is dry everywhere.  The initial surface is set so the wet
    # fraction matches the 2048 m workload's (35%).
    w0 = float(np.quantile(cell_bottom(B)[g:-g, g:-g], 0.35))
    U0 = terrain_initial_state(grid, B, w0)
    manning = 0.03
    xf, yf = grid.cell_faces_dir(0, interior=False), grid.cell_faces_dir(1, interior=False)
    # the target's bottom perturbation sits in the deepest water, away from dry ground
    depth = U0[:, :, 0] - cell_bottom(B)[g:-g, g:-g]
    ci, cj = np.unravel_index(np.argmax(depth), depth.shape)
    cx, cy = xf[ci + g] + 0.5, yf[cj + g] + 0.5
    bump = np.exp(-((xf[:, None] - cx) ** 2 + (yf[None, :] - cy) ** 2) / (2 * (0.04 * n) ** 2))

    def system_for(Bx, Cz):
        eq = VF.ShallowWaterEquations(VF.BottomTopography2D(Bx, None, grid), desingularizing_kappa=1e-4)
        return VF.ConservedSystem(None, VF.LinearReconstruction(), VF.CentralUpwind(eq), eq, grid,
                                  [VF.SourceTermBottom(), VF.ConstantRain(1e-5 * 3600)], VF.ImplicitFriction(Cz=Cz))

    def forward_J(Bx, Cz, target):
        sim = VF.Simulator(backend, system_for(Bx, Cz), VF.RungeKutta2(), grid, cfl=0.25)
        VF.set_current_state(sim, U0)
        VF.advance_steps(sim, steps)
        U = sim.get_full_state()
        h = U[g:-g, g:-g, 0] - cell_bottom(Bx)[g:-g, g:-g]
        return float(np.sum((h - target) ** 2)) if target is not None else h

    Bp = np.asfortranarray(B + 0.2 * bump)
    target = forward_J(Bp, manning ** 2, None)

    sim = VF.Simulator(backend, system_for(B, manning ** 2), VF.RungeKutta2(), grid, cfl=0.25)
    VF.set_current_state(sim, U0)
    t0 = time.perf_counter()
    VF.advance_steps(sim, steps)
    C.check(sim.handle, C.lib().sfv_synchronize(sim.handle))
    fwd_s = time.perf_counter() - t0
    VF.set_current_state(sim, U0)
    sim.t = 0.0
    VF.adjoint_gradient(VF.Simulator(backend, system_for(B, manning ** 2), VF.RungeKutta2(), grid), 2, target)  # warm-up
    t0 = time.perf_counter()
    res = VF.adjoint_gradient(sim, steps, target, checkpoint_every=checkpoint_every)
    adj_s = time.perf_counter() - t0
    gB, gn = res["dJ_dB"], VF.manning_gradient(res["dJ_dCz"], manning)

    rng = np.random.default_rng(seed)
    corner_depth = np.full(B.shape, -1.0)
    corner_depth[g + 1:-g - 1, g + 1:-g - 1] = np.minimum.reduce(
        [depth[:-1, :-1], depth[1:, :-1], depth[:-1, 1:], depth[1:, 1:]])   # all four cells of the corner
    wet = np.argwhere((bump > 0.3) & (corner_depth > 0.5))
    picks = [tuple(wet[k]) for k in rng.choice(len(wet), 3, replace=False)]
    checks = []
    eps = 1e-3
    for (i, j) in picks:
        Bq = B.copy(order="F"); Bq[i, j] += eps
        Bm = B.copy(order="F"); Bm[i, j] -= eps
        fd = (forward_J(Bq, manning ** 2, target) - forward_J(Bm, manning ** 2, target)) / (2 * eps)
        checks.append({"corner": [int(i), int(j)], "adjoint": float(gB[i, j]), "fd": fd,
                       "rel": abs(gB[i, j] - fd) / max(abs(fd), 1e-300)})
    hn = 1e-3 * manning
    fdn = (forward_J(B, (manning + hn) ** 2, target) - forward_J(B, (manning - hn) ** 2, target)) / (2 * hn)
    out = {
        "workload": (f"{n}x{n} terrain (seeded bumps, rain, Manning n = {manning}, initial surface "
                     f"w = max(B, {w0:.3f} m) for 35% wet cells), f64 RK2, {steps} steps, "
                     f"checkpoints every {checkpoint_every}; J = sum (h_T - h_target)^2, h_target from a run on "
                     f"B + 0.2 m Gaussian bump; gradients for all {gB.size} bottom corners and n"),
        "J": res["J"], "forward_s": fwd_s, "adjoint_s": adj_s, "adjoint_over_forward": adj_s / fwd_s,
        "dJ_dn": gn,
        "fd_corners": checks,
        "fd_dJ_dn": {"fd": fdn, "step": hn,
                     "note": ("J(n) jumps wherever a cell's depth crosses the dry cut-off (the clamp zeroes its "
                              "momentum), so central differences in n straddle jumps and do not converge; the "
                              "oracle's forward mode below is the check for dJ/dn")},
    }
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # the reference restatement, used only to check results
        dB = np.asfortranarray(rng.normal(size=B.shape) * bump)
        cfg = VF.build_config(system_for(B, manning ** 2), VF.RungeKutta2(), grid, 0.25, C.SFV_F64)
        t1 = time.perf_counter()
        _, (jvp, jvc) = O.ad_run2(cfg, B, U0, steps, target, (dB, 0.0, None), (None, 1.0, None))
        got = float(np.sum(gB * dB))
        jvn = VF.manning_gradient(jvc, manning)
        out["oracle_forward_mode"] = {"random_dB": {"adjoint": got, "oracle": jvp,
                                                    "rel": abs(got - jvp) / max(abs(jvp), 1e-300)},
                                      "dJ_dn": {"adjoint": gn, "oracle": jvn, "rel": abs(gn - jvn) / max(abs(jvn), 1e-300)},
                                      "oracle_s": time.perf_counter() - t1}
    except (ImportError, OSError) as e:      # the oracle library is not built
        out["oracle_forward_mode"] = {"skipped": str(e)}
    return out


def variant_rain(t, x, y):
    """A smooth per-cell rain field for the FunctionalRain variant (m/s)."""
    return 1e-5 * (1.0 + 0.5 * np.sin(x / 97.0) * np.cos(y / 61.0)) * (1.0 + 1e-3 * t)


def variants_bench(size: int, seed: int, steps: int, check: bool) -> dict:
    """The other configurations of the step on the terrain workload (same
    grid, bottom, initial state and friction), each timed device-resident on
    the GPU: ForwardEuler and NeumannBC (the hybrid kernel path), PeriodicBC and
    NoReconstruction (the march kernels), FunctionalRain (the host evaluates
    the rain function on every cell and uploads the field before each step,
    so its rate includes that work), plus the 1D Stoker dam break (1024
    cells to t = 20 s).  With `check`, each first step (1D: the whole run) is
    compared with the oracle's."""
    grid0, _, _, U0, B = build_scenario(size, "f64", seed)
    backend = VF.make_hip_backend(np.float64)
    L = C.lib()
    if check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # the reference restatement, used only to check results
        O.set_threads(os.cpu_count() or 1)

    def system_for(bc, rec, rain_fn):
        g = VF.CartesianGrid(size, size, gc=2, boundary=bc, extent=[[0.0, float(size)], [0.0, float(size)]])
        eq = VF.ShallowWaterEquations(VF.BottomTopography2D(B, None, g), desingularizing_kappa=1e-4)
        rain = VF.FunctionalRain(rain_fn, g) if rain_fn else VF.ConstantRain(1e-5 * 3600)
        system = VF.ConservedSystem(None, rec, VF.CentralUpwind(eq), eq, g, [VF.SourceTermBottom(), rain],
                                    VF.ImplicitFriction(Cz=0.03 * 0.03))
        return g, system, rain

    cases = [
        ("forward_euler", VF.WallBC(), VF.LinearReconstruction(), VF.ForwardEulerStepper(), None),
        ("neumann_bc", VF.NeumannBC(), VF.LinearReconstruction(), VF.RungeKutta2(), None),
        ("periodic_bc", VF.PeriodicBC(), VF.LinearReconstruction(), VF.RungeKutta2(), None),
        ("no_reconstruction", VF.WallBC(), VF.NoReconstruction(), VF.RungeKutta2(), None),
        ("functional_rain", VF.WallBC(), VF.LinearReconstruction(), VF.RungeKutta2(), variant_rain),
    ]
    out = {}
    for name, bc, rec, stepper, rain_fn in cases:
        g, system, rain = system_for(bc, rec, rain_fn)
        sim = VF.Simulator(backend, system, stepper, g, cfl=0.25)
        VF.set_current_state(sim, U0)
        VF.perform_step(sim, float("inf"))
        first_gpu = sim.get_full_state()
        VF.set_current_state(sim, U0)
        sim.t = 0.0
        VF.advance_steps(sim, 5)
        C.check(sim.handle, L.sfv_synchronize(sim.handle))
        t0 = time.perf_counter()
        VF.advance_steps(sim, steps)
        C.check(sim.handle, L.sfv_synchronize(sim.handle))
        el = time.perf_counter() - t0
        row = {"value": g.nx * g.ny * steps / el, "unit": "cell-updates/s", "ms_per_step": el / steps * 1e3,
               "steps": steps, "kernels": sim.kernels,
               "state_finite": bool(np.isfinite(sim.get_full_state()).all())}
        if check:
            cfg = VF.build_config(system, stepper, g, 0.25, C.SFV_F64)
            o = O.Oracle2D(cfg, B)
            o.set_interior(U0)
            if rain_fn:
                o.set_rain_field(VF.rain_field_values(rain, g, 0.0))
            o.step(float("inf"), 0.0)
            row["parity_first_step"] = "bitwise" if np.array_equal(first_gpu, o.state()) else "MISMATCH"
        out[name] = row

    # the 1D workload: Stoker dam break, 1024 cells on [0, 1000] m, RK2, CFL 0.25, to t = 20 s
    g1 = VF.CartesianGrid(1024, gc=2, boundary=VF.WallBC(), extent=[[0.0, 1000.0]])
    eq1 = VF.ShallowWaterEquations1D(VF.ConstantBottomTopography())
    s1 = VF.ConservedSystem(None, VF.LinearReconstruction(), VF.CentralUpwind(eq1), eq1, g1, [], None)
    U1 = np.zeros((1024, 2))
    U1[:, 0] = np.where(g1.cell_centers() < 500.0, 2.0, 1.0)
    sim = VF.Simulator(backend, s1, VF.RungeKutta2(), g1, cfl=0.25)
    VF.set_current_state(sim, U1)
    VF.simulate_to_time(sim, 20.0)            # warm-up run (tuning, first launches)
    VF.set_current_state(sim, U1)
    t1, n1 = ctypes.c_double(0.0), ctypes.c_int64()
    C.check(sim.handle, L.sfv_synchronize(sim.handle))
    t0 = time.perf_counter()
    C.check(sim.handle, L.sfv_run(sim.handle, 20.0, 1, 0.0, 0, ctypes.byref(t1), ctypes.byref(n1)))
    el = time.perf_counter() - t0
    n1 = n1.value
    row = {"value": 1024 * n1 / el, "unit": "cell-updates/s", "steps": int(n1), "seconds": el,
           "ms_per_step": el / max(n1, 1) * 1e3, "t_end": t1.value, "kernels": sim.kernels}
    if check:
        g2, s2 = VF.embed_1d(s1, g1)
        cfg = VF.build_config(s2, VF.RungeKutta2(), g2, 0.25, C.SFV_F64)
        cfg.ny = 1
        cfg.dy = cfg.dx
        o = O.Oracle1D(cfg, np.ascontiguousarray(s1.equation.B.corners(g1), dtype=np.float64))
        o.set_interior(U1)
        o.simulate_to_time(20.0)
        row["parity_full_run"] = "bitwise" if np.array_equal(sim.get_full_state(), o.state()) else "MISMATCH"
    out["stoker_1d"] = row
    return out


def grid_leg(size: int, precision: str, seed: int, steps: int, warmup: int, kernels: str = "march") -> dict:
    """One terrain run on the GPU: the first step kept for the parity probe,
    then `warmup` untimed and `steps` timed device-resident steps, and each
    kernel's mean device time from HIP events inside full steps."""
    grid, system, realtype, U0, B = build_scenario(size, precision, seed)
    backend = VF.make_hip_backend(realtype)
    sim = VF.Simulator(backend, system, VF.RungeKutta2(), grid, cfl=0.25)
    sim.set_kernels(kernels)
    L = C.lib()

    # parity probe: the first step, compared bitwise with the oracle's
    VF.set_current_state(sim, U0)
    VF.perform_step(sim, float("inf"))
    first_gpu = sim.get_full_state()

    VF.set_current_state(sim, U0)
    sim.t = 0.0
    if warmup > 0:
        VF.advance_steps(sim, warmup)
    C.check(sim.handle, L.sfv_synchronize(sim.handle))
    t0 = time.perf_counter()
    VF.advance_steps(sim, steps)
    C.check(sim.handle, L.sfv_synchronize(sim.handle))
    elapsed = time.perf_counter() - t0
    value = grid.nx * grid.ny * steps / elapsed
    finite = bool(np.isfinite(sim.get_full_state()).all())

    ms = {}
    timing_sim = sim
    if kernels == "per-pass":       # the stage kernels timed are the fused march ones
        timing_sim = VF.Simulator(backend, system, VF.RungeKutta2(), grid, cfl=0.25)
        VF.set_current_state(timing_sim, U0)
    for k in ("stage1", "stage2", "bc", "step", "copy"):
        out = ctypes.c_double()
        C.check(timing_sim.handle, L.sfv_time_kernel(timing_sim.handle, k.encode(), 20, ctypes.byref(out)))
        ms[k] = out.value
    elem = 8 if precision == "f64" else 4
    copy_bytes = 2 * 3 * (grid.nx + 2 * grid.gc) * (grid.ny + 2 * grid.gc) * elem   # read + write of one state
    return {"scenario": (grid, system, realtype, U0, B), "value": value, "elapsed": elapsed, "ms": ms,
            "march_rows": sim.march_rows,
            "dominant": max(("stage1", "stage2"), key=lambda k: ms[k]), "first_gpu": first_gpu,
            "finite": finite, "copy_gbs": copy_bytes / (ms["copy"] * 1e-3) / 1e9}


def f32_leg(args) -> dict:
    """The same terrain workload with Float32 storage (make_cpu_backend(Float32)
    semantics: f64 arithmetic rounded to f32 on every store, B and dt f64):
    throughput, kernel times and the first step compared with the oracle."""
    leg = grid_leg(args.size, "f32", args.seed, args.steps, args.warmup, args.kernels)
    grid, system, realtype, U0, B = leg["scenario"]
    cells = grid.nx * grid.ny
    dom = leg["dominant"]
    achieved = BYTES_PER_CELL_STAGE["f32"][dom] * cells / (leg["ms"][dom] * 1e-3) / 1e9
    out = {"value": leg["value"], "unit": "cell-updates/s", "ms_per_step": leg["elapsed"] / args.steps * 1e3,
           "steps": args.steps, "kernel_ms": leg["ms"],
           "roofline": {"kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS,
                        "algorithmic_bytes_per_launch": BYTES_PER_CELL_STAGE["f32"][dom] * cells},
           "state_finite": leg["finite"]}
    if not args.no_cpu:
        first_cpu = oracle_first_step(grid, system, realtype, U0, B)
        out["parity_first_step"] = "bitwise" if np.array_equal(leg["first_gpu"], first_cpu) else "MISMATCH"
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--precision", choices=["f64", "f32"], default="f64")
    ap.add_argument("--seed", type=int, default=2048)
    ap.add_argument("--kernels", choices=["march", "tiled", "per-pass"], default="march",
                    help="kernel variant of the RK2 step (all bitwise identical)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=48, help="steps per timed CPU run")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--ensemble", type=int, default=256, help="ensemble members (0: skip the ensemble leg)")
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--serving-rate", type=float, default=50.0, help="arrivals per second (0: skip serving)")
    ap.add_argument("--serving-duration", type=float, default=1.0, help="seconds of arrivals")
    ap.add_argument("--serving-tend", type=float, nargs=2, default=[1.0, 30.0], help="t_end range of a request (s)")
    ap.add_argument("--serving-budget-gb", type=float, default=64.0)
    ap.add_argument("--serving-cap", type=float, default=0.0,
                    help="stop each serving path this many seconds after the first arrival (0: run to completion)")
    ap.add_argument("--no-f32", dest="f32_leg", action="store_false",
                    help="skip the Float32-storage run of the same workload")
    ap.add_argument("--adjoint-size", type=int, default=512, help="adjoint leg grid side (0: skip)")
    ap.add_argument("--variant-steps", type=int, default=50,
                    help="timed steps of each configuration variant (0: skip the variants leg)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs on one GPU; --gpus must be 1 (multi-GPU sharding is a "
                                   "design note, not built)", "n_gpus": args.gpus}))
        return 2
    rank = int(os.environ.get("RANK", "0"))

    leg = grid_leg(args.size, args.precision, args.seed, args.steps, args.warmup, args.kernels)
    grid, system, realtype, U0, B = leg["scenario"]
    value, elapsed, ms, dominant = leg["value"], leg["elapsed"], leg["ms"], leg["dominant"]
    cells = grid.nx * grid.ny
    alg_bytes = BYTES_PER_CELL_STAGE[args.precision][dominant] * cells
    achieved = alg_bytes / (ms[dominant] * 1e-3) / 1e9
    workload = f"terrain{args.size}_{args.precision}"
    traffic = read_traffic(dominant, workload)
    step_gbs = value * BYTES_PER_CELL_STEP[args.precision] / 1e9
    first_gpu, finite, copy_gbs = leg["first_gpu"], leg["finite"], leg["copy_gbs"]
    # the FP64 VALU side: committed SQ_INSTS_VALU per launch against this run's kernel times
    valu = {k: read_valu(k, workload) for k in ("stage1", "stage2")}
    valu_frac = lane_cell = lane_update = None
    if valu[dominant] is not None:
        valu_frac = (valu[dominant] / VALU_WAVE_INSTR_PER_S) / (ms[dominant] * 1e-3)
        lane_cell = valu[dominant] * 64 / cells
    if None not in valu.values():
        lane_update = (valu["stage1"] + valu["stage2"]) * 64 / cells

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {
            "workload": (f"terrain {args.size}x{args.size}, dx=1 m, 32 seeded Gaussian bumps + 1e-3 slope, "
                         f"w=max(B,2 m), Manning n=0.03, rain 1e-5 m/s, wall BCs, kappa=1e-4, RK2, CFL 0.25, "
                         f"{args.steps} steps"),
            "nx": args.size, "ny": args.size, "precision": args.precision,
            "kernels": args.kernels,
            "march_rows": leg["march_rows"],
            "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm",
            "binding": "valu_f64" if valu_frac is not None and valu_frac > achieved / HBM_PEAK_GBS else "hbm",
            "valu_issue_frac": valu_frac,
            "lane_instr_per_cell": lane_cell,
            "lane_instr_per_cell_update": lane_update,
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": alg_bytes,
            "kernel_ms": ms,
            "copy_kernel_gbs": copy_gbs,
        },
        "step_roofline": {"bytes_per_cell_update": BYTES_PER_CELL_STEP[args.precision],
                          "achieved_gbs": step_gbs, "frac": step_gbs / HBM_PEAK_GBS},
        "state_finite": finite,
    }

    if rank == 0 and not args.no_cpu:
        threads = os.cpu_count() or 1
        try:
            threads = min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))
        except ValueError:
            pass
        first_cpu, cpu_value, cpu_med = cpu_baseline(grid, system, realtype, U0, B, args.cpu_steps, args.cpu_runs,
                                                     threads)
        _, cpu1_value, _ = cpu_baseline(grid, system, realtype, U0, B, 2, 1, 1)
        nproc, cpu_model = host_cpu()
        result["cpu_baseline"] = {
            "value": cpu_value, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": cpu_model,
            "sample": (f"C++ oracle (loop-faithful restatement of the reference CPU path), same workload: "
                       f"median of {args.cpu_runs} runs of {args.cpu_steps} RK2 steps after 1 warm-up step, "
                       f"OpenMP over rows"),
            "single_thread_value": cpu1_value,
            "stoker_1d": cpu_stoker_1d(threads),
        }
        result["parity_first_step"] = "bitwise" if np.array_equal(first_gpu, first_cpu) else "MISMATCH"
        result["gpu_over_cpu"] = value / cpu_value
    if rank == 0 and args.precision == "f64" and args.f32_leg:
        result["f32"] = f32_leg(args)
    if rank == 0 and args.ensemble > 0:
        result["ensemble"] = ensemble_bench(args.ensemble, args.ensemble_steps, args.seed)
    if rank == 0 and args.serving_rate > 0:
        result["serving"] = serving_bench(args.serving_rate, args.serving_duration, tuple(args.serving_tend),
                                          args.serving_budget_gb, args.seed, args.serving_cap)
    if rank == 0 and args.adjoint_size > 0:
        result["adjoint"] = adjoint_bench(args.seed, args.adjoint_size)
    if rank == 0 and args.variant_steps > 0:
        result["variants"] = variants_bench(args.size, args.seed, args.variant_steps, not args.no_cpu)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
