"""Simulator, HIP backend and the driver functions (src/simulator.jl,
src/backends/kernel_abstractions.jl, src/callbacks.jl).

The HIP backend is the only backend this package has: every step runs on the
gfx950 kernels of libsinfvm_hip.so.  If the library or a GPU is missing,
make_hip_backend() raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import warnings
from collections import namedtuple

import numpy as np

from . import _capi as C
from .grid import CartesianGrid
from .model import (EMBED_ROWS, ConservedSystem, FunctionalRain, RungeKutta2, build_config, embed_1d,
                    rain_field_values)

InteriorState = namedtuple("InteriorState", ["h", "hu", "hv"])
InteriorState1D = namedtuple("InteriorState1D", ["h", "hu"])


class HIPBackend:
    """Counterpart of KernelAbstractionBackend{KA, RealType}
    (kernel_abstractions.jl:29-34).  realtype float64 is the reference's
    Float64 backend; float32 is make_cpu_backend(Float32): Float32 storage
    with the reference's Float64 parameters."""

    def __init__(self, realtype=np.float64):
        self.realtype = np.dtype(realtype)
        if self.realtype not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise C.ArgumentError(C.SFV_ERR_ARGUMENT, "realtype must be float64 or float32")
        self.precision = C.SFV_F64 if self.realtype == np.float64 else C.SFV_F32_STORAGE


def device_count() -> int:
    return int(C.lib().sfv_device_count())


def make_hip_backend(realtype=np.float64) -> HIPBackend:
    """Load the HIP library and check a GPU is visible; raises otherwise."""
    if device_count() < 1:
        raise RuntimeError("no HIP device visible: the HIP backend needs an MI355X (gfx950)")
    return HIPBackend(realtype)


def has_hip_backend() -> bool:
    try:
        make_hip_backend()
        return True
    except Exception:
        return False


def get_available_backends():
    """kernel_abstractions.jl:45-55, for this package: the HIP backend if usable."""
    return [make_hip_backend()] if has_hip_backend() else []


def name(backend) -> str:
    return "HIP"


class Simulator:
    """Simulator(backend, system, timestepper, grid; cfl=0.25, t0=0.0)
    (simulator.jl:60-91).  Owns one device handle: state volumes, RHS,
    topography and the step-control block live in HBM."""

    def __init__(self, backend: HIPBackend, system: ConservedSystem, timestepper, grid: CartesianGrid, *,
                 cfl: float = 0.25, t0: float = 0.0):
        self.backend = backend
        self.system = system
        self.timestepper = timestepper
        self.grid = grid
        self.realtype = backend.realtype
        # cfl, t and current_timestep are stored as realtype (simulator.jl:40-44)
        self.cfl = float(self.realtype.type(cfl))
        self.t = float(self.realtype.type(t0))
        self.current_timestep = 0.0
        # A CartesianGrid{1} run is the exactly equivalent 2D run on (nx, 3)
        # (model.embed_1d); the state the user sees is its middle row.
        self._grid2, self._system2 = (embed_1d(system, grid) if grid.dimension == 1 else (grid, system))
        self._cfg = build_config(self._system2, timestepper, self._grid2, cfl, backend.precision)
        L = C.lib()
        h = ctypes.c_void_p()
        C.check(None, L.sfv_create(ctypes.byref(self._cfg), ctypes.byref(h)))
        self._h = h
        B = np.asfortranarray(self._system2.equation.B.corners(self._grid2), dtype=np.float64)
        C.check(self._h, L.sfv_set_topography(self._h, C.dptr(B), B.shape[0], B.shape[1]))
        fr = [s for s in self._system2.source_terms if isinstance(s, FunctionalRain)]
        self._functional_rain = fr[0] if fr else None
        self._rain_buf = None
        if self._functional_rain is not None:
            from ._pinned import PinnedArray
            self._rain_buf = PinnedArray(self._grid2.totalcells)   # page-locked: the upload runs at DMA speed

    def _refresh_rain(self) -> None:
        """FunctionalRain: the rates at this step's start time on every cell,
        evaluated by host threads into the page-locked buffer and uploaded."""
        if self._functional_rain is not None:
            F = rain_field_values(self._functional_rain, self._grid2, self.t, out=self._rain_buf.array)
            C.check(self._h, C.lib().sfv_set_rain_field(self._h, C.dptr(F)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            C.lib().sfv_destroy(h)
            self._h = None
        buf = getattr(self, "_rain_buf", None)
        if buf is not None:
            buf.free()
            self._rain_buf = None

    # -- state transfer ----------------------------------------------------
    @property
    def handle(self):
        return self._h

    def _state_shape(self):
        if self.grid.dimension == 1:
            return (self.grid.totalcells[0], 2)
        return (*self.grid.totalcells, 3)

    def get_full_state(self) -> np.ndarray:
        """The full state with ghost cells: (nxT, nyT, 3), or (nxT, 2) in 1D."""
        U = np.empty((*self._grid2.totalcells, 3), dtype=self.realtype, order="F")
        C.check(self._h, C.lib().sfv_get_state(self._h, U.ctypes.data_as(ctypes.c_void_p), C.SFV_LAYOUT_FULL))
        if self.grid.dimension == 1:
            return np.array(U[:, self.grid.gc + 1, :2], order="F")
        return U

    def set_full_state(self, U) -> None:
        U = np.asfortranarray(U, dtype=self.realtype)
        if U.shape != self._state_shape():
            raise C.DomainError(C.SFV_ERR_DOMAIN, f"state should be of size {self._state_shape()} but got {U.shape}")
        if self.grid.dimension == 1:
            U2 = np.zeros((*self._grid2.totalcells, 3), dtype=self.realtype, order="F")
            U2[:, :, :2] = U[:, None, :]
            U = U2
        C.check(self._h, C.lib().sfv_set_state(self._h, U.ctypes.data_as(ctypes.c_void_p), C.SFV_LAYOUT_FULL))

    KERNELS = {"per-pass": 0, "tiled": 1, "march": 2, "hybrid": 3}

    def set_kernels(self, variant: str) -> None:
        """Choose the kernels of a step: "march" (the RK2 default), "tiled",
        "hybrid" (march stage-1 right-hand sides with per-pass updates: the
        default for ForwardEuler and NeumannBC) or "per-pass".  All give
        bitwise identical results."""
        C.check(self._h, C.lib().sfv_set_fused(self._h, self.KERNELS[variant]))

    @property
    def kernels(self) -> str:
        """The kernels a step runs: "march", "tiled", "hybrid" or "per-pass"."""
        v = C.lib().sfv_get_fused(self._h)
        return {k: n for n, k in self.KERNELS.items()}[v]

    def set_fused(self, fused: bool) -> None:
        self.set_kernels("march" if fused else "per-pass")

    def set_march_rows(self, rows: int) -> None:
        """Rows per wave segment of the march kernels: 0 (default) = measured
        on the first step and cached per grid size class; > 0 forces it.
        Results are bitwise the same for every value."""
        C.check(self._h, C.lib().sfv_set_march_rows(self._h, int(rows)))

    @property
    def march_rows(self) -> int:
        """The segment height in use (0 until the first step chooses it)."""
        return int(C.lib().sfv_get_march_rows(self._h))

    def _step_status(self, rc: int) -> None:
        if rc == C.SFV_ERR_NONFINITE_DT:
            # the reference carries on with the NaN/Inf dt (simulator.jl:190-192)
            warnings.warn("a time step produced a non-finite dt", RuntimeWarning, stacklevel=3)
            return
        C.check(self._h, rc)


def current_state(sim: Simulator) -> np.ndarray:
    """collect(current_state(sim)): the full (nxT, nyT, 3) state, ghosts included."""
    return sim.get_full_state()


def current_interior_state(sim: Simulator) -> InteriorState:
    """current_interior_state(sim) (simulator.jl:113), per-variable interior views."""
    U = sim.get_full_state()
    g = sim.grid.gc
    if sim.grid.dimension == 1:
        I = U[g:-g, :]
        return InteriorState1D(np.array(I[:, 0]), np.array(I[:, 1]))
    I = U[g:-g, g:-g, :]
    return InteriorState(np.array(I[:, :, 0]), np.array(I[:, :, 1]), np.array(I[:, :, 2]))


def current_time(sim: Simulator) -> float:
    return sim.t


def current_timestep(sim: Simulator) -> float:
    return sim.current_timestep


def _as_interior_array(sim: Simulator, new_state) -> np.ndarray:
    if sim.grid.dimension == 1:
        nx = sim.grid.nx
        if isinstance(new_state, (tuple, list)) and len(new_state) == 2:
            arr = np.stack([np.asarray(a, dtype=np.float64) for a in new_state], axis=-1)
        else:
            arr = np.asarray(new_state, dtype=np.float64)
        if arr.shape != (nx, 2):
            raise C.DomainError(C.SFV_ERR_DOMAIN, f"state should be of size {(nx, 2)} but got {arr.shape}")
        U2 = np.zeros((nx, EMBED_ROWS, 3))
        U2[:, :, :2] = arr[:, None, :]
        return np.asfortranarray(U2.astype(sim.realtype))
    nx, ny = sim.grid.interior_size()
    if isinstance(new_state, (tuple, list)) and len(new_state) == 3:
        arr = np.stack([np.asarray(a, dtype=np.float64) for a in new_state], axis=-1)
    else:
        arr = np.asarray(new_state, dtype=np.float64)
    if arr.shape != (nx, ny, 3):
        raise C.DomainError(C.SFV_ERR_DOMAIN, f"state should be of size {(nx, ny, 3)} but got {arr.shape}")
    return np.asfortranarray(arr.astype(sim.realtype))


def set_current_state(sim: Simulator, new_state) -> None:
    """set_current_state!(sim, state) (simulator.jl:119-135): write the interior
    (nx, ny, 3) — or a tuple (w, hu, hv) of (nx, ny) arrays — and apply the
    boundary condition; corner ghost cells keep their values."""
    U = _as_interior_array(sim, new_state)
    C.check(sim.handle, C.lib().sfv_set_state(sim.handle, U.ctypes.data_as(ctypes.c_void_p), C.SFV_LAYOUT_INTERIOR))


def perform_step(sim: Simulator, max_dt: float) -> float:
    """perform_step!(sim, max_dt) (simulator.jl:145-177): one step with the
    sources at sim.t.  Like the reference it does not advance sim.t."""
    dt = ctypes.c_double()
    sim._refresh_rain()
    sim._step_status(C.lib().sfv_step(sim.handle, float(max_dt), sim.t, ctypes.byref(dt)))
    sim.current_timestep = dt.value
    return dt.value


def _realtype_add(sim: Simulator, a: float, b: float) -> float:
    T = sim.realtype.type
    return float(T(T(a) + T(b)))


def simulate_to_time(sim: Simulator, endtime: float, *, match_endtime: bool = True, callback=None,
                     show_progress: bool = False, maximum_timestep=None) -> None:
    """simulate_to_time(sim, endtime; match_endtime, callback, show_progress,
    maximum_timestep) (simulator.jl:179-212).  Without a callback the whole
    loop runs device-resident (no host synchronisation per step); with a
    callback it runs step by step on the host and calls callback(t, sim)
    after every step, as the reference does."""
    if callback is not None and not callable(callback):
        callback = MultipleCallbacks(callback)
    L = C.lib()
    if callback is None and sim._functional_rain is None:
        t = ctypes.c_double(sim.t)
        steps = ctypes.c_int64()
        mts = float(maximum_timestep) if maximum_timestep is not None else 0.0
        rc = L.sfv_run(sim.handle, float(endtime), 1 if match_endtime else 0, mts, 0, ctypes.byref(t),
                       ctypes.byref(steps))
        sim.t = t.value
        if steps.value > 0:
            dt = ctypes.c_double()
            C.check(sim.handle, L.sfv_current_timestep(sim.handle, ctypes.byref(dt)))
            sim.current_timestep = dt.value
        sim._step_status(rc)
        return
    while sim.t < endtime:
        max_dt = endtime - sim.t if match_endtime else float("inf")
        if maximum_timestep is not None:
            max_dt = min(max_dt, maximum_timestep)
        perform_step(sim, max_dt)
        sim.t = _realtype_add(sim, sim.t, sim.current_timestep)
        if callback is not None:
            callback(sim.t, sim)


def advance_steps(sim: Simulator, nsteps: int, max_dt: float = float("inf")) -> None:
    """nsteps of perform_step!(sim, max_dt) with t += dt, device-resident
    (host-driven per step when a FunctionalRain needs its rates each step)."""
    if sim._functional_rain is not None:
        for _ in range(int(nsteps)):
            perform_step(sim, max_dt)
            sim.t = _realtype_add(sim, sim.t, sim.current_timestep)
        return
    t = ctypes.c_double(sim.t)
    sim._step_status(C.lib().sfv_advance(sim.handle, int(nsteps), float(max_dt), ctypes.byref(t)))
    sim.t = t.value
    dt = ctypes.c_double()
    C.check(sim.handle, C.lib().sfv_current_timestep(sim.handle, ctypes.byref(dt)))
    sim.current_timestep = dt.value


class IntervalWriter:
    """IntervalWriter(step, writer) (callbacks.jl:40-56): calls writer(t, sim)
    when t + dt reaches the next output time."""

    def __init__(self, step: float = 1.0, writer=None, current_t: float = 0.0):
        if writer is None and callable(step):
            step, writer = 1.0, step
        self.step = float(step)
        self.writer = writer
        self.current_t = float(current_t)

    def __call__(self, t, sim):
        dt = current_timestep(sim)
        if t + dt >= self.current_t:
            self.writer(t, sim)
            self.current_t += self.step


class MultipleCallbacks:
    """MultipleCallbacks(callbacks) (callbacks.jl:61-82)."""

    def __init__(self, callbacks):
        self.callbacks = list(callbacks)

    def __call__(self, t, sim):
        for c in self.callbacks:
            c(t, sim)
