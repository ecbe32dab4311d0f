"""C ABI and host-side checks that need no GPU: the library loads and exports
every entry point include/sinfvm_hip.h declares, configuration validation
returns the status codes that mirror the reference's exceptions, the Python
mirror raises the same exception types as the reference constructors, and
the product path never reaches the CPU oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

import volumefluxes as VF
from volumefluxes import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sinfvm_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*[\w\s\*]+?\b(sfv_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = C.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def _valid_config():
    cfg = C.SfvConfig()
    C.lib().sfv_config_defaults(ctypes.byref(cfg), 16, 8, 1.0, 1.0)
    return cfg


def _create(cfg):
    h = ctypes.c_void_p()
    rc = C.lib().sfv_create(ctypes.byref(cfg), ctypes.byref(h))
    if rc == C.SFV_OK:
        C.lib().sfv_destroy(h)
    return rc, C.lib().sfv_create_error().decode()


def test_config_defaults_mirror_reference():
    cfg = _valid_config()
    assert (cfg.gc, cfg.bc, cfg.g, cfg.rho) == (2, C.SFV_BC_WALL, 9.81, 1.0)
    assert cfg.depth_cutoff == 1e-5 and cfg.desingularizing_kappa == 1e-5     # swe_2D.jl:28
    assert cfg.theta == 1.2 and cfg.cfl == 0.25                               # linearreconstruction.jl:24, simulator.jl:72
    assert cfg.Cz == 0.03 * 0.03                                              # friction.jl:26
    assert (cfg.horton_f0, cfg.horton_fc, cfg.horton_k) == (1.977e-4, 3.272e-5, 2.43e-3)


@pytest.mark.parametrize("field,value,code", [
    ("nx", 0, C.SFV_ERR_DOMAIN),
    ("nx", (1 << 27) - 16, C.SFV_ERR_DOMAIN),   # rows are addressed by 31-bit byte offsets
    ("gc", 0, C.SFV_ERR_DOMAIN),
    ("gc", 1, C.SFV_ERR_ASSERTION),          # LinearReconstruction needs gc > 1 (linearreconstruction.jl:59)
    ("bc", 7, C.SFV_ERR_ARGUMENT),
    ("stepper", 3, C.SFV_ERR_ARGUMENT),
    ("friction", 9, C.SFV_ERR_ARGUMENT),
    ("precision", 5, C.SFV_ERR_ARGUMENT),
])
def test_create_validation_errors(field, value, code):
    cfg = _valid_config()
    setattr(cfg, field, value)
    rc, msg = _create(cfg)
    assert rc == code and msg


def test_time_table_without_entries_is_domain_error():
    cfg = _valid_config()
    cfg.rain = C.SFV_RAIN_TIME_TABLE
    cfg.rain_table_len = 0
    assert _create(cfg)[0] == C.SFV_ERR_DOMAIN


def test_device_bytes_for_2048_f64():
    cfg = _valid_config()
    cfg.nx = cfg.ny = 2048
    b = C.lib().sfv_device_bytes_for(ctypes.byref(cfg))
    nxT = 2052
    # ping-pong pair + right-hand side (the fused kernels need no stage-1 state)
    assert b >= 3 * 3 * nxT * nxT * 8 + 2053 * 2053 * 8
    assert b < 3 * 3 * nxT * nxT * 8 + 2053 * 2053 * 8 + 1e6
    cfg.bc = C.SFV_BC_NEUMANN            # per-pass path: + the stage-1 state
    assert C.lib().sfv_device_bytes_for(ctypes.byref(cfg)) >= 4 * 3 * nxT * nxT * 8


def test_batch_create_argument_errors():
    L = C.lib()
    b = ctypes.c_void_p()
    assert L.sfv_batch_create(None, 0, ctypes.byref(b)) == C.SFV_ERR_ARGUMENT
    assert L.sfv_batch_last_error(None)
    arr = (ctypes.c_void_p * 2)(None, None)
    assert L.sfv_batch_create(arr, 2, ctypes.byref(b)) == C.SFV_ERR_ARGUMENT
    assert not b.value
    assert L.sfv_batch_size(None) == 0


# ---- the Python mirror raises the reference's exception types
def test_nonzero_bottom_without_source_is_argument_error():
    grid = VF.CartesianGrid(4, 4, gc=2)
    B = np.ones(tuple(t + 1 for t in grid.size))
    eq = VF.ShallowWaterEquations(VF.BottomTopography2D(B, None, grid))
    with pytest.raises(VF.ArgumentError):                # equation.jl:74-84
        VF.ConservedSystem(None, VF.LinearReconstruction(), VF.CentralUpwind(eq), eq, grid, [])
    VF.ConservedSystem(None, VF.LinearReconstruction(), VF.CentralUpwind(eq), eq, grid, [VF.SourceTermBottom()])


def test_rain_table_size_mismatch_is_domain_error():
    with pytest.raises(VF.DomainError):                  # source_term_rain.jl:36-38
        VF.TimeDependentRain([1.0, 2.0], [0.0])


def test_infiltration_factor_size_is_domain_error():
    grid = VF.CartesianGrid(10, 10, gc=2)
    with pytest.raises(VF.DomainError):                  # source_term_infiltration.jl:61-65
        VF.HortonInfiltration(VF.CartesianGrid(2, 2), None, factor=np.ones((10, 10)))
    assert VF.HortonInfiltration(grid).factor.shape == grid.size


def test_functional_rain_maps_to_a_rate_field():
    grid = VF.CartesianGrid(4, 4, gc=2)
    eq = VF.ShallowWaterEquations()
    system = VF.ConservedSystem(None, VF.LinearReconstruction(), VF.CentralUpwind(eq), eq, grid,
                                [VF.FunctionalRain(lambda t, x, y: 0.0, grid)])
    cfg = VF.build_config(system, VF.RungeKutta2(), grid, 0.25, C.SFV_F64)
    assert cfg.rain == C.SFV_RAIN_FIELD


def test_build_config_maps_reference_system():
    grid = VF.CartesianGrid(32, 16, gc=2, boundary=VF.PeriodicBC(), extent=[[0.0, 64.0], [0.0, 8.0]])
    eq = VF.ShallowWaterEquations(depth_cutoff=1e-6, desingularizing_kappa=1e-4)
    system = VF.ConservedSystem(None, VF.LinearReconstruction(1.3), VF.CentralUpwind(eq), eq, grid,
                                [VF.HortonInfiltration(grid), VF.ConstantRain(0.036)],
                                VF.ImplicitFriction(Cz=0.05 ** 2, friction_function=VF.friction_fcg2016))
    cfg = VF.build_config(system, VF.ForwardEulerStepper(), grid, 0.2, C.SFV_F64)
    assert (cfg.nx, cfg.ny, cfg.dx, cfg.dy) == (32, 16, 2.0, 0.5)
    assert cfg.bc == C.SFV_BC_PERIODIC and cfg.theta == 1.3 and cfg.stepper == C.SFV_FORWARD_EULER
    assert cfg.friction == C.SFV_FRICTION_FCG2016 and cfg.Cz == 0.05 ** 2
    assert cfg.rain == C.SFV_RAIN_CONSTANT and cfg.infiltration == C.SFV_INFILTRATION_HORTON
    assert cfg.rain_before_infiltration == 0          # infiltration precedes rain in the list
    assert cfg.bottom_source == 0 and cfg.depth_cutoff == 1e-6


def test_hip_backend_fails_loudly_without_gpu():
    if C.lib().sfv_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        VF.make_hip_backend()
    assert VF.get_available_backends() == []


def test_product_code_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "volumefluxes.jl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert not re.search(r"^\s*(import|from)\s+oracle\b", text, flags=re.M), f
                assert "liboracle" not in text and "swe_oracle" not in text, f


def test_server_needs_a_device():
    L = C.lib()
    s = ctypes.c_void_p()
    assert L.sfv_server_create(0, 0, ctypes.byref(s)) == C.SFV_ERR_ARGUMENT
    if L.sfv_device_count() > 0:
        pytest.skip("a GPU is visible")
    assert L.sfv_server_create(1 << 30, 0, ctypes.byref(s)) == C.SFV_ERR_HIP
    assert not s.value
    with pytest.raises(VF.SfvError if hasattr(VF, "SfvError") else Exception):
        VF.Server(1 << 30)
