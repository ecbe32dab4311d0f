"""Gradients of a shallow-water run with respect to its bottom topography,
friction coefficient and initial state: the reverse-mode adjoint of the
time stepper on the GPU (sfv_adjoint in the C ABI).

The reference differentiates runs with ForwardDiff (examples/optimization.jl),
one forward pass per parameter direction; the adjoint gives the whole
gradient (every bottom corner, Cz and the initial state) for about the cost
of a few forward runs, with checkpoints bounding the memory.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C


def adjoint_gradient(sim, nsteps: int, h_target, *, checkpoint_every: int = 20, want_state: bool = True) -> dict:
    """J = sum((h_T - h_target)^2) over interior cells after nsteps of
    perform_step!(sim, Inf) from sim's current state, and its gradient:
    {"J", "dJ_dB" (corners, like B), "dJ_dCz", "dJ_dU0" (interior (nx, ny, 3))}.
    The simulator ends at the final state and time."""
    g = sim.grid
    if g.dimension != 2:
        raise C.ArgumentError(C.SFV_ERR_ARGUMENT, "the adjoint runs 2D grids")
    tgt = np.asfortranarray(h_target, dtype=np.float64)
    if tgt.shape != (g.nx, g.ny):
        raise C.DomainError(C.SFV_ERR_DOMAIN, f"h_target should be of size {(g.nx, g.ny)} but got {tgt.shape}")
    dB = np.zeros((g.totalcells[0] + 1, g.totalcells[1] + 1), order="F")
    dU = np.zeros((g.nx, g.ny, 3), order="F") if want_state else None
    J = ctypes.c_double()
    dCz = ctypes.c_double()
    t = ctypes.c_double(sim.t)
    C.check(sim.handle, C.lib().sfv_adjoint(sim.handle, int(nsteps), int(checkpoint_every), ctypes.byref(t),
                                            C.dptr(tgt), ctypes.byref(J), C.dptr(dB), ctypes.byref(dCz),
                                            C.dptr(dU)))
    sim.t = t.value
    dt = ctypes.c_double()
    C.check(sim.handle, C.lib().sfv_current_timestep(sim.handle, ctypes.byref(dt)))
    sim.current_timestep = dt.value
    return {"J": J.value, "dJ_dB": dB, "dJ_dCz": dCz.value, "dJ_dU0": dU}


def manning_gradient(dJ_dCz: float, n: float) -> float:
    """dJ/dn for Cz = n^2 (friction.jl:26 uses Cz = n^2 for Manning's n)."""
    return 2.0 * n * dJ_dCz
