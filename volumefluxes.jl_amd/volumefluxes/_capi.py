"""ctypes view of the C ABI declared in include/sinfvm_hip.h.

Loads ``libsinfvm_hip.so`` (the HIP backend for gfx950) from the package's
in-tree build directory.  There is no CPU fallback: if the library is missing
or cannot be loaded, :func:`lib` raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int32, c_int64, c_void_p

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
NATIVE_DIR = os.path.join(os.path.dirname(PKG_DIR), "lib")
LIB_NAME = "libsinfvm_hip.so"

# ---- enums (values must match include/sinfvm_hip.h)
SFV_OK = 0
SFV_ERR_DOMAIN = 1
SFV_ERR_ARGUMENT = 2
SFV_ERR_ASSERTION = 3
SFV_ERR_OUT_OF_MEMORY = 4
SFV_ERR_NONFINITE_DT = 5
SFV_ERR_HIP = 6
SFV_ERR_STATE = 7
SFV_ERR_BUDGET = 8

SFV_BC_WALL, SFV_BC_PERIODIC, SFV_BC_NEUMANN = 0, 1, 2
SFV_FORWARD_EULER, SFV_RUNGE_KUTTA2 = 1, 2
SFV_NO_RECONSTRUCTION, SFV_LINEAR_RECONSTRUCTION = 0, 1
SFV_FRICTION_NONE, SFV_FRICTION_BH2021, SFV_FRICTION_BSA2012, SFV_FRICTION_FCG2016 = 0, 1, 2, 3
SFV_RAIN_NONE, SFV_RAIN_CONSTANT, SFV_RAIN_TIME_TABLE, SFV_RAIN_FIELD = 0, 1, 2, 3
SFV_INFILTRATION_NONE, SFV_INFILTRATION_CONSTANT, SFV_INFILTRATION_HORTON = 0, 1, 2
SFV_F64, SFV_F32_STORAGE = 0, 1
SFV_LAYOUT_FULL, SFV_LAYOUT_INTERIOR = 0, 1


class SfvConfig(ctypes.Structure):
    """Mirror of ``struct sfv_config``."""

    _fields_ = [
        ("nx", c_int32), ("ny", c_int32), ("gc", c_int32),
        ("dx", c_double), ("dy", c_double),
        ("bc", c_int32),
        ("g", c_double), ("rho", c_double), ("depth_cutoff", c_double),
        ("desingularizing_kappa", c_double),
        ("reconstruction", c_int32), ("theta", c_double),
        ("stepper", c_int32), ("cfl", c_double),
        ("bottom_source", c_int32),
        ("friction", c_int32), ("Cz", c_double),
        ("rain", c_int32), ("rain_rate", c_double),
        ("rain_table_len", c_int32),
        ("rain_table_rates", POINTER(c_double)), ("rain_table_times", POINTER(c_double)),
        ("infiltration", c_int32), ("infiltration_rate", c_double),
        ("horton_f0", c_double), ("horton_fc", c_double), ("horton_k", c_double),
        ("infiltration_factor_scalar", c_double),
        ("infiltration_factor", POINTER(c_double)),
        ("rain_before_infiltration", c_int32),
        ("precision", c_int32),
        ("embedded_1d", c_int32),
    ]

    def keepalive(self, *arrays):
        """Hold numpy arrays whose pointers this struct carries."""
        self.__dict__.setdefault("_keep", []).extend(arrays)


class SfvError(RuntimeError):
    """An error status returned by the C ABI."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[sfv status {code}] {message}")
        self.code = code


class DomainError(SfvError, ValueError):
    """Mirror of Julia's DomainError (array size mismatches)."""


class ArgumentError(SfvError, ValueError):
    """Mirror of Julia's ArgumentError."""


class SfvAssertionError(SfvError, AssertionError):
    """Mirror of Julia's AssertionError (e.g. gc < 2 with linear reconstruction)."""


class NonFiniteTimestepError(SfvError, FloatingPointError):
    """dt came out NaN or infinite (e.g. an all-dry domain)."""


class BudgetError(SfvError, MemoryError):
    """A serving request does not fit the HBM byte budget."""


_ERROR_TYPES = {
    SFV_ERR_DOMAIN: DomainError,
    SFV_ERR_ARGUMENT: ArgumentError,
    SFV_ERR_ASSERTION: SfvAssertionError,
    SFV_ERR_NONFINITE_DT: NonFiniteTimestepError,
    SFV_ERR_BUDGET: BudgetError,
}


class RequestInfo(ctypes.Structure):
    """Mirror of ``struct sfv_request_info``."""

    _fields_ = [("id", c_int64), ("status", c_int32), ("error", c_int32), ("t", c_double), ("steps", c_int64),
                ("cell_updates", c_int64), ("device_bytes", c_int64), ("submit_s", c_double),
                ("admit_s", c_double), ("done_s", c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(ctypes.Structure):
    """Mirror of ``struct sfv_server_stats``."""

    _fields_ = [(k, c_int64) for k in ("submitted", "admitted", "completed", "failed", "queued", "running",
                                       "bytes_in_use", "bytes_peak", "budget_bytes", "cell_updates", "replays")] + \
               [("busy_s", c_double), ("uptime_s", c_double), ("host_admit_s", c_double), ("host_retire_s", c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


SFV_REQ_QUEUED, SFV_REQ_RUNNING, SFV_REQ_DONE, SFV_REQ_FAILED = 0, 1, 2, 3


def check_server(server, code: int) -> None:
    if code != SFV_OK:
        msg = lib().sfv_server_last_error(server)
        raise_for(code, (msg or b"").decode(errors="replace"))


def raise_for(code: int, message: str) -> None:
    if code == SFV_OK:
        return
    raise _ERROR_TYPES.get(code, SfvError)(code, message)


_LIB = None


def lib_path() -> str:
    return os.environ.get("SINFVM_HIP_LIB", os.path.join(NATIVE_DIR, LIB_NAME))


def _declare(L) -> None:
    cfgp = POINTER(SfvConfig)
    h = c_void_p
    L.sfv_config_defaults.argtypes = [cfgp, c_int32, c_int32, c_double, c_double]
    L.sfv_config_defaults.restype = None
    L.sfv_create.argtypes = [cfgp, POINTER(c_void_p)]
    L.sfv_create.restype = c_int32
    L.sfv_destroy.argtypes = [h]
    L.sfv_destroy.restype = None
    L.sfv_last_error.argtypes = [h]
    L.sfv_last_error.restype = c_char_p
    L.sfv_create_error.argtypes = []
    L.sfv_create_error.restype = c_char_p
    L.sfv_set_topography.argtypes = [h, POINTER(c_double), c_int64, c_int64]
    L.sfv_set_topography.restype = c_int32
    L.sfv_set_state.argtypes = [h, c_void_p, c_int32]
    L.sfv_set_rain_field.argtypes = [h, POINTER(c_double)]
    L.sfv_set_rain_field.restype = c_int32
    L.sfv_adjoint.argtypes = [h, c_int64, c_int32, POINTER(c_double), POINTER(c_double), POINTER(c_double),
                              POINTER(c_double), POINTER(c_double), POINTER(c_double)]
    L.sfv_adjoint.restype = c_int32
    L.sfv_set_state.restype = c_int32
    L.sfv_get_state.argtypes = [h, c_void_p, c_int32]
    L.sfv_get_state.restype = c_int32
    L.sfv_step.argtypes = [h, c_double, c_double, POINTER(c_double)]
    L.sfv_step.restype = c_int32
    L.sfv_run.argtypes = [h, c_double, c_int32, c_double, c_int64, POINTER(c_double), POINTER(c_int64)]
    L.sfv_run.restype = c_int32
    L.sfv_advance.argtypes = [h, c_int64, c_double, POINTER(c_double)]
    L.sfv_advance.restype = c_int32
    L.sfv_advance_noalloc.argtypes = [cfgp, c_void_p, c_void_p, c_int64, c_double, POINTER(c_double),
                                      POINTER(c_double), c_void_p]
    L.sfv_advance_noalloc.restype = c_int32
    L.sfv_current_timestep.argtypes = [h, POINTER(c_double)]
    L.sfv_current_timestep.restype = c_int32
    L.sfv_synchronize.argtypes = [h]
    L.sfv_synchronize.restype = c_int32
    L.sfv_device_bytes.argtypes = [h]
    L.sfv_device_bytes.restype = c_int64
    L.sfv_device_bytes_for.argtypes = [cfgp]
    L.sfv_device_bytes_for.restype = c_int64
    L.sfv_time_kernel.argtypes = [h, c_char_p, c_int32, POINTER(c_double)]
    L.sfv_time_kernel.restype = c_int32
    L.sfv_device_count.argtypes = []
    L.sfv_device_count.restype = c_int32
    L.sfv_set_fused.argtypes = [h, c_int32]
    L.sfv_set_fused.restype = c_int32
    L.sfv_get_fused.argtypes = [h]
    L.sfv_get_fused.restype = c_int32
    L.sfv_set_march_rows.argtypes = [h, c_int32]
    L.sfv_set_march_rows.restype = c_int32
    L.sfv_get_march_rows.argtypes = [h]
    L.sfv_get_march_rows.restype = c_int32
    b = c_void_p
    L.sfv_batch_create.argtypes = [POINTER(c_void_p), c_int32, POINTER(c_void_p)]
    L.sfv_batch_create.restype = c_int32
    L.sfv_batch_destroy.argtypes = [b]
    L.sfv_batch_destroy.restype = None
    L.sfv_batch_last_error.argtypes = [b]
    L.sfv_batch_last_error.restype = c_char_p
    L.sfv_batch_size.argtypes = [b]
    L.sfv_batch_size.restype = c_int32
    L.sfv_batch_run.argtypes = [b, POINTER(c_double), c_int32, c_double, c_int64, POINTER(c_double),
                                POINTER(c_int64)]
    L.sfv_batch_run.restype = c_int32
    L.sfv_batch_get_march_rows.argtypes = [b, POINTER(c_int32)]
    L.sfv_batch_get_march_rows.restype = c_int32
    sv = c_void_p
    L.sfv_server_create.argtypes = [c_int64, c_int32, POINTER(c_void_p)]
    L.sfv_server_create.restype = c_int32
    L.sfv_server_destroy.argtypes = [sv]
    L.sfv_server_destroy.restype = None
    L.sfv_server_last_error.argtypes = [sv]
    L.sfv_server_last_error.restype = c_char_p
    L.sfv_server_submit.argtypes = [sv, cfgp, POINTER(c_double), c_int64, c_int64, c_void_p, c_double, c_double,
                                    c_int32, c_double, POINTER(c_int64)]
    L.sfv_server_submit.restype = c_int32
    L.sfv_server_poll.argtypes = [sv, c_int64, POINTER(RequestInfo)]
    L.sfv_server_poll.restype = c_int32
    L.sfv_server_wait.argtypes = [sv, c_int64, c_double, POINTER(RequestInfo)]
    L.sfv_server_wait.restype = c_int32
    L.sfv_server_fetch.argtypes = [sv, c_int64, c_void_p, POINTER(RequestInfo)]
    L.sfv_server_fetch.restype = c_int32
    L.sfv_server_get_stats.argtypes = [sv, POINTER(ServerStats)]
    L.sfv_server_get_stats.restype = c_int32


def lib():
    """Load libsinfvm_hip.so (once).  Raises RuntimeError if it is absent."""
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"HIP backend library not found at {path}; build it with "
                "`python __graft_entry__.py build` (hipcc --offload-arch=gfx950)")
        _LIB = ctypes.CDLL(path)
        _declare(_LIB)
    return _LIB


def check(handle, code: int) -> None:
    if code != SFV_OK:
        msg = lib().sfv_last_error(handle) if handle else lib().sfv_create_error()
        raise_for(code, (msg or b"").decode(errors="replace"))


def check_batch(batch, code: int) -> None:
    if code != SFV_OK:
        msg = lib().sfv_batch_last_error(batch)
        raise_for(code, (msg or b"").decode(errors="replace"))


def dptr(arr):
    """POINTER(c_double) to a contiguous float64 numpy array (or NULL)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(POINTER(c_double))
