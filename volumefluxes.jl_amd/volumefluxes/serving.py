"""Serving: flood scenarios submitted at any time, run on one GPU with
continuous batching under an HBM byte budget.

No counterpart in the reference.  A Server owns a native worker thread
(sfv_server_* in the C ABI) that admits requests in arrival order while
their device bytes fit the budget, advances every admitted request together
with the batched kernels (each with its own dt, to its own end time, as
simulate_to_time does, simulator.jl:179-212), and retires finished ones:
results are copied to host memory and their device memory freed.  A result
is bitwise that of running the request alone.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C
from .model import ConservedSystem, RungeKutta2, build_config


class Server:
    """Server(budget_bytes, realtype=np.float64)"""

    def __init__(self, budget_bytes: int, realtype=np.float64):
        self.realtype = np.dtype(realtype).type
        self.precision = C.SFV_F32_STORAGE if self.realtype == np.float32 else C.SFV_F64
        s = ctypes.c_void_p()
        C.check_server(None, C.lib().sfv_server_create(int(budget_bytes), self.precision, ctypes.byref(s)))
        self._s = s
        self._shapes = {}

    def close(self) -> None:
        s = getattr(self, "_s", None)
        if s:
            C.lib().sfv_server_destroy(s)
            self._s = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, system: ConservedSystem, grid, U0, t_end: float, *, t0: float = 0.0, cfl: float = 0.25,
               stepper=None, match_endtime: bool = True, maximum_timestep=None) -> int:
        """Queue one scenario: run from t0 to t_end from the interior state U0
        (nx, ny, 3).  Returns the request id."""
        stepper = stepper if stepper is not None else RungeKutta2()
        cfg = build_config(system, stepper, grid, cfl, self.precision)
        B = np.asfortranarray(system.equation.B.corners(grid), dtype=np.float64)
        U = np.asfortranarray(U0, dtype=self.realtype)
        if U.shape != (grid.nx, grid.ny, 3):
            raise C.DomainError(C.SFV_ERR_DOMAIN, f"state should be of size {(grid.nx, grid.ny, 3)} but got {U.shape}")
        rid = ctypes.c_int64()
        mts = float(maximum_timestep) if maximum_timestep is not None else 0.0
        C.check_server(self._s, C.lib().sfv_server_submit(
            self._s, ctypes.byref(cfg), C.dptr(B), B.shape[0], B.shape[1], U.ctypes.data_as(ctypes.c_void_p),
            float(t0), float(t_end), 1 if match_endtime else 0, mts, ctypes.byref(rid)))
        self._shapes[rid.value] = (grid.nx, grid.ny, 3)
        return rid.value

    def poll(self, rid: int) -> dict:
        info = C.RequestInfo()
        C.check_server(self._s, C.lib().sfv_server_poll(self._s, int(rid), ctypes.byref(info)))
        return info.as_dict()

    def wait(self, rid: int, timeout: float | None = None) -> dict:
        """Block until the request has finished (or failed); TimeoutError after
        `timeout` seconds."""
        info = C.RequestInfo()
        rc = C.lib().sfv_server_wait(self._s, int(rid), -1.0 if timeout is None else float(timeout),
                                     ctypes.byref(info))
        if rc == C.SFV_ERR_STATE:
            raise TimeoutError(f"request {rid} not finished after {timeout} s")
        C.check_server(self._s, rc)
        return info.as_dict()

    def fetch(self, rid: int):
        """(final interior state, info) of a finished request, which is then
        forgotten.  A failed request raises its error (BudgetError, ...)."""
        shape = self._shapes.get(int(rid))
        if shape is None:
            raise C.ArgumentError(C.SFV_ERR_ARGUMENT, f"unknown request {rid}")
        U = np.empty(shape, dtype=self.realtype, order="F")
        info = C.RequestInfo()
        rc = C.lib().sfv_server_fetch(self._s, int(rid), U.ctypes.data_as(ctypes.c_void_p), ctypes.byref(info))
        if rc == C.SFV_ERR_STATE and info.status in (C.SFV_REQ_QUEUED, C.SFV_REQ_RUNNING):
            C.check_server(self._s, rc)          # not finished: the request stays
        self._shapes.pop(int(rid), None)
        if rc == C.SFV_ERR_NONFINITE_DT:
            import warnings
            warnings.warn((C.lib().sfv_server_last_error(self._s) or b"").decode(), RuntimeWarning)
        else:
            C.check_server(self._s, rc)
        return U, info.as_dict()

    def stats(self) -> dict:
        st = C.ServerStats()
        C.check_server(self._s, C.lib().sfv_server_get_stats(self._s, ctypes.byref(st)))
        return st.as_dict()
