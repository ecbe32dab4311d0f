"""Ensembles: several Simulators advanced together on one GPU.

No counterpart in the reference, whose user steps one Simulator at a time
(simulator.jl:179-212).  An Ensemble groups Simulators of any grid size,
each with its own system, topography, state, time and time step, and
advances them with one persistent kernel launch per RK stage for the whole
group (a captured HIP graph replayed on the device).  Each member's results
are bitwise those of advancing it alone with simulate_to_time /
advance_steps.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _capi as C
from .simulator import Simulator


class Ensemble:
    """Ensemble(sims): a batch over the Simulators' device handles.  The
    Simulators stay usable (get_full_state, current_state, ...) between runs;
    keep them alive while the Ensemble exists."""

    def __init__(self, sims: Sequence[Simulator]):
        self.sims = list(sims)
        if not self.sims:
            raise C.ArgumentError(C.SFV_ERR_ARGUMENT, "an ensemble needs at least one simulator")
        n = len(self.sims)
        arr = (ctypes.c_void_p * n)(*[s.handle for s in self.sims])
        b = ctypes.c_void_p()
        C.check_batch(None, C.lib().sfv_batch_create(arr, n, ctypes.byref(b)))
        self._b = b

    def __len__(self) -> int:
        return len(self.sims)

    @property
    def march_rows(self) -> int:
        """Segment height the last run used for members without a forced one
        (timed against its neighbours on the run's first replays; 0 before
        any run).  It never changes results."""
        r = ctypes.c_int32()
        C.check_batch(self._b, C.lib().sfv_batch_get_march_rows(self._b, ctypes.byref(r)))
        return int(r.value)

    def close(self) -> None:
        b = getattr(self, "_b", None)
        if b:
            C.lib().sfv_batch_destroy(b)
            self._b = None

    def __del__(self):
        self.close()

    def _run(self, t_end, match_endtime: bool, maximum_timestep, max_steps: int) -> np.ndarray:
        n = len(self.sims)
        t = (ctypes.c_double * n)(*[s.t for s in self.sims])
        steps = (ctypes.c_int64 * n)()
        te = None
        if t_end is not None:
            te_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(t_end, dtype=np.float64), (n,)))
            te = te_arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        mts = float(maximum_timestep) if maximum_timestep is not None else 0.0
        rc = C.lib().sfv_batch_run(self._b, te, 1 if match_endtime else 0, mts, int(max_steps), t, steps)
        for i, s in enumerate(self.sims):
            s.t = t[i]
            dt = ctypes.c_double()
            C.check(s.handle, C.lib().sfv_current_timestep(s.handle, ctypes.byref(dt)))
            s.current_timestep = dt.value
        if rc == C.SFV_ERR_NONFINITE_DT:
            import warnings
            warnings.warn((C.lib().sfv_batch_last_error(self._b) or b"").decode(), RuntimeWarning)
        else:
            C.check_batch(self._b, rc)
        return np.frombuffer(steps, dtype=np.int64).copy()

    def advance_steps(self, nsteps: int, max_dt: float = float("inf")) -> np.ndarray:
        """nsteps of perform_step!(sim, max_dt) with t += dt for every member
        (the batched form of advance_steps).  Returns the steps per member."""
        if nsteps <= 0:
            return np.zeros(len(self.sims), dtype=np.int64)
        mts = None if not np.isfinite(max_dt) else float(max_dt)
        return self._run(None, False, mts, int(nsteps))

    def simulate_to_time(self, endtime, *, match_endtime: bool = True, maximum_timestep=None,
                         max_steps: int = 0) -> np.ndarray:
        """simulate_to_time(sim, endtime[i]) for every member at once
        (simulator.jl:179-212, without callbacks).  endtime is a scalar or
        one value per member.  Returns the steps per member."""
        return self._run(endtime, match_endtime, maximum_timestep, int(max_steps))
