"""Wave timeline of the march stage kernels on the 2048^2 terrain workload.

Needs the measurement build of the library (the march kernels record each
wave's start and end, hardware slot and row start times):

    make -C volumefluxes.jl_amd variant V=timeline FLAGS=-DSFV_MARCH_TIMELINE=1
    SINFVM_HIP_LIB=volumefluxes.jl_amd/lib/variants/libsinfvm_hip_timeline.so SFV_AUTOTUNE=0 \\
        python3 volumefluxes.jl_amd/timeline_probe.py [n] [rows]

It steps the terrain 30 times, then records single steps and prints, per
stage: the launch span, wave durations, how full the resident wave slots are
over the launch (3 per SIMD), how many full rows the device finishes per 5 us,
and how far apart the waves sharing a SIMD end.  Timestamps are the 100 MHz
real-time counter (10 ns).
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(HERE), HERE]

import volumefluxes as VF  # noqa: E402
from volumefluxes import _capi as C  # noqa: E402
from volumefluxes.scenarios import terrain_workload  # noqa: E402

TL_WAVES, TL_ROWS, SLOTS_PER_SIMD = 16384, 32, 3


def record(sim, L):
    waves = np.zeros((2, TL_WAVES, 4), dtype=np.uint64)
    rows = np.zeros((2, TL_WAVES, TL_ROWS), dtype=np.uint64)
    if L.sfv_probe_timeline_clear() != 0:
        raise RuntimeError("clearing the timeline failed")
    VF.advance_steps(sim, 1)
    C.check(sim.handle, L.sfv_synchronize(sim.handle))
    if L.sfv_probe_timeline(waves.ctypes.data_as(ctypes.c_void_p), rows.ctypes.data_as(ctypes.c_void_p)) != 0:
        raise RuntimeError("reading the timeline failed")
    return waves, rows


def analyse(waves, rows, stage, my, ny):
    w = waves[stage]
    used = w[:, 1] > 0
    w, r = w[used], rows[stage][used]
    t0 = w[:, 0].astype(np.int64)
    base = t0.min()
    t0 -= base
    t1 = w[:, 1].astype(np.int64) - base
    span = int(t1.max())
    hw = w[:, 2].astype(np.int64)
    xcc = (w[:, 3] >> 32).astype(np.int64) & 0xF
    seg = (w[:, 3] & 0xFFFF).astype(np.int64)
    simd = ((((xcc * 8 + ((hw >> 13) & 7)) * 2 + ((hw >> 12) & 1)) * 16 + ((hw >> 8) & 15)) * 4 + ((hw >> 4) & 3))
    nsimd = len(np.unique(simd))
    slots = nsimd * SLOTS_PER_SIMD
    nit = my + 4
    full = (seg * my + my <= ny) & (r[:, min(nit, TL_ROWS) - 1] > 0)
    dur = (t1 - t0)[full] / 100.0
    ev = np.zeros(span + 2, dtype=np.int64)
    np.add.at(ev, t0, 1)
    np.add.at(ev, t1, -1)
    res = np.cumsum(ev)[:span + 1]
    occ = [np.mean(res[int(f * span):int((f + 0.05) * span) + 1]) / slots for f in np.arange(0, 1, 0.05)]
    R = r[full][:, :min(nit, TL_ROWS)].astype(np.int64) - base
    ends = np.concatenate([R[:, 4:], t1[full][:, None]], axis=1).ravel()   # ends of iterations 3 ..
    cnt, _ = np.histogram(ends, bins=np.arange(0, span + 500, 500))
    per_simd = {}
    for k, e in zip(simd[full], t1[full]):
        per_simd.setdefault(k, []).append(e)
    spread = [(max(v) - min(v)) / 100.0 for v in per_simd.values() if len(v) > 1]
    out = [f"stage {stage + 1}: launch span {span / 100:.1f} us, {len(w)} waves ({int(full.sum())} full-height) "
           f"on {nsimd} SIMDs ({slots} resident slots at {SLOTS_PER_SIMD} per SIMD)",
           f"  wave duration us: p1 {np.percentile(dur, 1):.1f}  p10 {np.percentile(dur, 10):.1f}  "
           f"median {np.median(dur):.1f}  p90 {np.percentile(dur, 90):.1f}  max {dur.max():.1f}",
           "  resident slots filled, by 5% of the span: " + " ".join(f"{x:.2f}" for x in occ),
           "  full rows finished per 5 us (device-wide): " + " ".join(str(int(c)) for c in cnt),
           f"  waves on one SIMD end within: median {np.median(spread):.1f} us, p90 {np.percentile(spread, 90):.1f} us"]
    return "\n".join(out)


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 2048
    rows_forced = int(sys.argv[2]) if len(sys.argv) > 2 else 0
    L = C.lib()
    if not hasattr(L, "sfv_probe_timeline"):
        raise RuntimeError("this library is not the timeline build (make variant V=timeline "
                           "FLAGS=-DSFV_MARCH_TIMELINE=1, then SINFVM_HIP_LIB=...)")
    L.sfv_probe_timeline.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    grid, system, realtype, U0, B = terrain_workload(n, "f64", 2048)
    sim = VF.Simulator(VF.HIPBackend(np.float64), system, VF.RungeKutta2(), grid, cfl=0.25)
    if rows_forced:
        sim.set_march_rows(rows_forced)
    VF.set_current_state(sim, U0)
    VF.advance_steps(sim, 30)
    C.check(sim.handle, L.sfv_synchronize(sim.handle))
    snaps = [record(sim, L) for _ in range(3)]
    my = sim.march_rows
    print(f"terrain {n}x{n} f64 RK2, {my}-row segments, step 34 of the run (the last of three recorded)")
    waves, rows = snaps[-1]
    for stage in (0, 1):
        print(analyse(waves, rows, stage, my, n))


if __name__ == "__main__":
    main()
