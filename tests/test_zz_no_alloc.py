"""sfv_advance_noalloc (the entry point that allocates nothing) against
sfv_advance on a handle, through the C ABI: two calls in a row on one GPU,
the second with a smaller grid (the static scratch keeps the first call's
contents), then Float32 storage and a periodic grid; each result (full state
with ghosts, t and dt) bitwise equal to the handle's.  Device buffers for the
inputs come from hipMalloc through ctypes; the entry point itself allocates
nothing."""
import ctypes
import ctypes.util

import numpy as np
import pytest

import volumefluxes as VF
from volumefluxes import _capi as C

from helpers import make_hip, terrain

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            lib = ctypes.CDLL(name)
            break
        except OSError:
            lib = None
    assert lib is not None, "libamdhip64.so not found"
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


class DeviceArray:
    """A hipMalloc'd copy of a host array (test plumbing only)."""

    def __init__(self, hip, a):
        self.hip, self.shape, self.dtype = hip, a.shape, a.dtype
        self.nbytes = a.nbytes
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        a = np.asfortranarray(a)
        assert hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _same(a, b):
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na], b[~nb])


def _check(hip, sc, nsteps, max_dt=float("inf")):
    # the handle path (existing entry point): its full initial state is the input
    sim = make_hip(sc)
    U0 = sim.get_full_state()
    VF.advance_steps(sim, nsteps, max_dt)
    want, t_want, dt_want = sim.get_full_state(), sim.t, sim.current_timestep
    cfg = sc.config()
    B = DeviceArray(hip, sc.corners())
    U = DeviceArray(hip, U0)
    try:
        t, dt = ctypes.c_double(0.0), ctypes.c_double(-1.0)
        rc = C.lib().sfv_advance_noalloc(ctypes.byref(cfg), B.ptr, U.ptr, nsteps, max_dt, ctypes.byref(t),
                                         ctypes.byref(dt), None)
        assert rc == C.SFV_OK, (rc, C.lib().sfv_create_error())
        got = U.get()
    finally:
        U.free()
        B.free()
    print(f"[noalloc] {sc.name} {got.dtype} {nsteps} steps: t={t.value!r} dt={dt.value!r}", flush=True)
    assert t.value == t_want and dt.value == dt_want
    assert got.dtype == want.dtype and _same(got, want)


def test_noalloc_matches_sfv_advance_twice_then_smaller():
    hip = _hip()
    _check(hip, terrain(300, 200, seed=5), 7)
    _check(hip, terrain(120, 90, seed=6), 4)          # smaller input: the scratch holds the first call's data


def test_noalloc_float32_periodic_and_max_dt():
    hip = _hip()
    _check(hip, terrain(150, 130, seed=7, realtype=np.float32), 5)
    _check(hip, terrain(90, 70, seed=8, bc=VF.PeriodicBC()), 6, max_dt=0.01)


def test_noalloc_rejects_what_it_cannot_run():
    hip = _hip()
    small = DeviceArray(hip, np.zeros(64))
    try:
        def call(cfg):
            t = ctypes.c_double(0.0)
            return C.lib().sfv_advance_noalloc(ctypes.byref(cfg), small.ptr, small.ptr, 1, float("inf"),
                                               ctypes.byref(t), None, None)
        big = terrain(1200, 1200, seed=1).config()              # 1204^2 cells > 1,398,085
        assert call(big) == C.SFV_ERR_DOMAIN
        assert b"exceeds" in C.lib().sfv_create_error()
        fe = terrain(40, 30, seed=1, stepper=VF.ForwardEulerStepper()).config()
        assert call(fe) == C.SFV_ERR_ARGUMENT
        neu = terrain(40, 30, seed=1, bc=VF.NeumannBC()).config()
        assert call(neu) == C.SFV_ERR_ARGUMENT
    finally:
        small.free()
