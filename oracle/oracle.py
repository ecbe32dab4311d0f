"""ctypes wrapper of the CPU ORACLE (liboracle_swe.so) — test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, to check the HIP backend against it and to time the CPU
baseline.  It mirrors the reference's Simulator for CartesianGrid{2} and {1}
(see swe_oracle.hpp for the per-function reference citations).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_swe.so")
_LIB = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        from volumefluxes._capi import SfvConfig
        cfgp = POINTER(SfvConfig)
        d = c_double
        L.oracle_set_threads.argtypes = [c_int]
        L.oracle_get_threads.restype = c_int
        for n in ("oracle_minmod",):
            getattr(L, n).argtypes = [d, d, d]
            getattr(L, n).restype = d
        for n in ("oracle_jl_max", "oracle_jl_min", "oracle_desingularize"):
            getattr(L, n).argtypes = [d, d]
            getattr(L, n).restype = d
        L.oracle_compute_rain.argtypes = [cfgp, d]
        L.oracle_compute_rain.restype = d
        L.oracle_compute_infiltration.argtypes = [cfgp, d, c_int64]
        L.oracle_compute_infiltration.restype = d
        L.oracle_central_upwind.argtypes = [cfgp, c_int, POINTER(d), POINTER(d), POINTER(d)]
        L.oracle_central_upwind.restype = d
        L.oracle2d_create.argtypes = [cfgp]
        L.oracle2d_create.restype = c_void_p
        L.oracle2d_destroy.argtypes = [c_void_p]
        L.oracle2d_set_topography.argtypes = [c_void_p, POINTER(d)]
        L.oracle2d_set_rain_field.argtypes = [c_void_p, POINTER(d)]
        L.oracle_ad_run.argtypes = [c_void_p, POINTER(d), POINTER(d), ctypes.c_int64, POINTER(d), POINTER(d), d,
                                    POINTER(d), POINTER(d), POINTER(d), POINTER(d)]
        L.oracle_ad_run.restype = d
        L.oracle_ad_run2.argtypes = [c_void_p, POINTER(d), POINTER(d), ctypes.c_int64, POINTER(d), POINTER(d), d,
                                     POINTER(d), POINTER(d), d, POINTER(d), POINTER(d)]
        L.oracle_ad_run2.restype = d
        L.oracle_ad_forward.argtypes = [c_void_p, POINTER(d), POINTER(d), ctypes.c_int64, POINTER(d), POINTER(d)]
        L.oracle_ad_forward.restype = None
        L.oracle2d_set_interior.argtypes = [c_void_p, POINTER(d)]
        L.oracle2d_set_full.argtypes = [c_void_p, c_void_p]
        L.oracle2d_get_full.argtypes = [c_void_p, c_void_p]
        L.oracle2d_get_substep.argtypes = [c_void_p, c_int, c_void_p]
        L.oracle2d_get_system_buffer.argtypes = [c_void_p, c_int, c_void_p]
        L.oracle2d_step.argtypes = [c_void_p, d, d]
        L.oracle2d_step.restype = d
        L.oracle2d_current_timestep.argtypes = [c_void_p]
        L.oracle2d_current_timestep.restype = d
        L.oracle2d_simulate_to_time.argtypes = [c_void_p, d, c_int, d, c_int64, POINTER(d)]
        L.oracle2d_simulate_to_time.restype = c_int64
        L.oracle2d_advance.argtypes = [c_void_p, c_int64, d, POINTER(d)]
        L.oracle2d_apply_bc.argtypes = [c_void_p]
        L.oracle2d_B_lookup.argtypes = [c_void_p, c_int, c_int, c_int, c_int]
        L.oracle2d_B_lookup.restype = d
        L.oracle1d_create.argtypes = [cfgp]
        L.oracle1d_create.restype = c_void_p
        L.oracle1d_destroy.argtypes = [c_void_p]
        L.oracle1d_set_topography.argtypes = [c_void_p, POINTER(d)]
        L.oracle1d_set_interior.argtypes = [c_void_p, POINTER(d)]
        L.oracle1d_get_full.argtypes = [c_void_p, POINTER(d)]
        L.oracle1d_step.argtypes = [c_void_p, d, d]
        L.oracle1d_step.restype = d
        L.oracle1d_simulate_to_time.argtypes = [c_void_p, d, c_int, d, c_int64, POINTER(d)]
        L.oracle1d_simulate_to_time.restype = c_int64
        L.oracle1d_current_timestep.argtypes = [c_void_p]
        L.oracle1d_current_timestep.restype = d
        _LIB = L
    return _LIB


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def _dp(a):
    return None if a is None else a.ctypes.data_as(POINTER(c_double))


class Oracle2D:
    """The reference's Simulator on a CartesianGrid{2}, on the CPU."""

    def __init__(self, cfg, B=None):
        self.cfg = cfg
        self.nx, self.ny, self.gc = cfg.nx, cfg.ny, cfg.gc
        self.nxT, self.nyT = self.nx + 2 * self.gc, self.ny + 2 * self.gc
        self.realtype = np.float32 if cfg.precision == 1 else np.float64
        self._h = lib().oracle2d_create(ctypes.byref(cfg))
        self.t = 0.0
        if B is not None:
            self.set_topography(B)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle2d_destroy(self._h)
            self._h = None

    def set_topography(self, B):
        B = np.asfortranarray(B, dtype=np.float64)
        assert B.shape == (self.nxT + 1, self.nyT + 1)
        self._B = B
        lib().oracle2d_set_topography(self._h, _dp(B))

    def set_rain_field(self, field) -> None:
        """FunctionalRain rates of the coming step on all (nxT, nyT) cells."""
        f = np.asfortranarray(field, dtype=np.float64)
        lib().oracle2d_set_rain_field(self._h, _dp(f))

    def B_lookup(self, which: int, i: int, j: int, d: int = 0) -> float:
        return lib().oracle2d_B_lookup(self._h, which, i, j, d)

    def apply_bc(self) -> None:
        lib().oracle2d_apply_bc(self._h)

    def set_interior(self, U):
        U = np.asfortranarray(U, dtype=np.float64)
        assert U.shape == (self.nx, self.ny, 3)
        lib().oracle2d_set_interior(self._h, _dp(U))

    def set_full(self, U):
        U = np.asfortranarray(U, dtype=self.realtype)
        assert U.shape == (self.nxT, self.nyT, 3)
        lib().oracle2d_set_full(self._h, U.ctypes.data_as(c_void_p))

    def state(self) -> np.ndarray:
        U = np.empty((self.nxT, self.nyT, 3), dtype=self.realtype, order="F")
        lib().oracle2d_get_full(self._h, U.ctypes.data_as(c_void_p))
        return U

    def substep(self, k: int) -> np.ndarray:
        U = np.empty((self.nxT, self.nyT, 3), dtype=self.realtype, order="F")
        lib().oracle2d_get_substep(self._h, k, U.ctypes.data_as(c_void_p))
        return U

    def system_buffer(self, which: int) -> np.ndarray:
        shape = (self.nxT, self.nyT) if which == 2 else (self.nxT, self.nyT, 3)
        U = np.empty(shape, dtype=self.realtype, order="F")
        lib().oracle2d_get_system_buffer(self._h, which, U.ctypes.data_as(c_void_p))
        return U

    def step(self, max_dt: float, t: float | None = None) -> float:
        return lib().oracle2d_step(self._h, float(max_dt), self.t if t is None else float(t))

    @property
    def current_timestep(self) -> float:
        return lib().oracle2d_current_timestep(self._h)

    def simulate_to_time(self, endtime, match_endtime=True, maximum_timestep=None, max_steps=0) -> int:
        t = c_double(self.t)
        n = lib().oracle2d_simulate_to_time(self._h, float(endtime), 1 if match_endtime else 0,
                                            float(maximum_timestep) if maximum_timestep else 0.0,
                                            int(max_steps), ctypes.byref(t))
        self.t = t.value
        return n

    def advance(self, nsteps: int, max_dt: float = float("inf")) -> None:
        t = c_double(self.t)
        lib().oracle2d_advance(self._h, int(nsteps), float(max_dt), ctypes.byref(t))
        self.t = t.value


class Oracle1D:
    """The reference's Simulator on a CartesianGrid{1} (ShallowWaterEquations1D)."""

    def __init__(self, cfg, B=None):
        self.cfg = cfg
        self.nx, self.gc = cfg.nx, cfg.gc
        self.nxT = self.nx + 2 * self.gc
        self._h = lib().oracle1d_create(ctypes.byref(cfg))
        self.t = 0.0
        if B is not None:
            B = np.ascontiguousarray(B, dtype=np.float64)
            assert B.shape == (self.nxT + 1,)
            self._B = B
            lib().oracle1d_set_topography(self._h, _dp(B))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle1d_destroy(self._h)
            self._h = None

    def set_interior(self, U):
        U = np.asfortranarray(U, dtype=np.float64)
        assert U.shape == (self.nx, 2)
        lib().oracle1d_set_interior(self._h, _dp(U))

    def state(self) -> np.ndarray:
        U = np.empty((self.nxT, 2), dtype=np.float64, order="F")
        lib().oracle1d_get_full(self._h, _dp(U))
        return U

    def step(self, max_dt: float, t: float | None = None) -> float:
        return lib().oracle1d_step(self._h, float(max_dt), self.t if t is None else float(t))

    @property
    def current_timestep(self) -> float:
        return lib().oracle1d_current_timestep(self._h)

    def simulate_to_time(self, endtime, match_endtime=True, maximum_timestep=None, max_steps=0) -> int:
        t = c_double(self.t)
        n = lib().oracle1d_simulate_to_time(self._h, float(endtime), 1 if match_endtime else 0,
                                            float(maximum_timestep) if maximum_timestep else 0.0,
                                            int(max_steps), ctypes.byref(t))
        self.t = t.value
        return n


# ---- forward-mode derivatives (oracle_ad.hpp)
def _f(a):
    return None if a is None else np.asfortranarray(a, dtype=np.float64)


def ad_run(cfg, B, U0, nsteps, target, dB=None, dCz=0.0, dU0=None):
    """J = sum (h_T - target)^2 after nsteps steps (max_dt = Inf) and its
    derivative along (dB, dCz, dU0).  Returns (J, dJ, U_T full, dts)."""
    B, U0, target, dB, dU0 = _f(B), _f(U0), _f(target), _f(dB), _f(dU0)
    nxT, nyT = cfg.nx + 2 * cfg.gc, cfg.ny + 2 * cfg.gc
    U = np.empty((nxT, nyT, 3), order="F")
    dts = np.empty(int(nsteps))
    dJ = c_double()
    J = lib().oracle_ad_run(ctypes.byref(cfg), _dp(B), _dp(U0), int(nsteps), _dp(target), _dp(dB), float(dCz),
                            _dp(dU0), ctypes.byref(dJ), _dp(U), _dp(dts))
    return J, dJ.value, U, dts


def ad_forward(cfg, B, U0, nsteps):
    """The same run on plain doubles: (U_T full, dts)."""
    B, U0 = _f(B), _f(U0)
    nxT, nyT = cfg.nx + 2 * cfg.gc, cfg.ny + 2 * cfg.gc
    U = np.empty((nxT, nyT, 3), order="F")
    dts = np.empty(int(nsteps))
    lib().oracle_ad_forward(ctypes.byref(cfg), _dp(B), _dp(U0), int(nsteps), _dp(U), _dp(dts))
    return U, dts


def ad_run2(cfg, B, U0, nsteps, target, dir1, dir2):
    """Two directional derivatives in one pass; dir = (dB, dCz, dU0).
    Returns (J, [dJ1, dJ2])."""
    B, U0, target = _f(B), _f(U0), _f(target)
    (b1, c1, u1), (b2, c2, u2) = dir1, dir2
    b1, u1, b2, u2 = _f(b1), _f(u1), _f(b2), _f(u2)
    out = (c_double * 2)()
    J = lib().oracle_ad_run2(ctypes.byref(cfg), _dp(B), _dp(U0), int(nsteps), _dp(target), _dp(b1), float(c1),
                             _dp(u1), _dp(b2), float(c2), _dp(u2), out)
    return J, [out[0], out[1]]
