"""Closed-form pins of the oracle's ImplicitFriction (friction.jl:30-84).

A uniform flow over a flat bottom with periodic boundaries has every
reconstruction slope, flux difference and bottom-slope term exactly zero, so a
step changes the state only through the implicit friction update.  Each
substep then reduces to the scalar recurrence the reference's
`implicit_friction` defines (friction.jl:62-78) applied with the substep's
INPUT state (simulator.jl:168-172: `implicit_substep!(outputs[s+1],
outputs[s], ...)`):

    ForwardEuler:  hu1 = hu0 / (1 - dt*f(h, hu0, hv0))
    RungeKutta2:   hu1 = hu0 / (1 - dt*f(h, hu0, hv0))
                   hu2 = 0.5*(hu0 + hu1) / (1 - dt*f(h, hu1, hv1))

with f = -Cz*|u|/h*^p, p = 2 (bh2021), 4/3 (bsa2012), 7/3 (fcg2016), and
h* the desingularised depth (equation.jl:56).  The expected values below are
computed independently of the oracle's code (numpy `hypot` and real powers
instead of sqrt(u*u+v*v), h*h and cbrt(h)*h), so they agree to a few ulps;
the tolerance is 1e-14 relative.  A slip such as /h instead of /h^2, friction
taken from the stage output, or the wrong sign on dt*f changes the result at
the 1e-3 level or more and fails these tests.
"""
import math

import numpy as np
import pytest

import oracle as O
from volumefluxes import _capi as C

POWER = {C.SFV_FRICTION_BH2021: 2.0, C.SFV_FRICTION_BSA2012: 4.0 / 3.0, C.SFV_FRICTION_FCG2016: 7.0 / 3.0}
RTOL = 1e-14


def _desing(h, kappa):
    # Brodtkorb & Holm 2021 eq. 26, written from the paper: h* = sign(h) max(|h|, min(h^2/(2κ) + κ/2, κ))
    return math.copysign(1.0, h) * max(abs(h), min(h * h / (2 * kappa) + kappa / 2, kappa))


def _factor(law, Cz, h, hu, hv, kappa):
    hs = _desing(h, kappa)
    s = float(np.hypot(hu / hs, hv / hs))
    return -Cz * s / hs ** POWER[law]


def _cfg(law, stepper, *, nx=6, ny=5, Cz=0.03 ** 2, kappa=1e-5, cutoff=1e-5, dx=1.0, dy=1.0):
    cfg = C.SfvConfig()
    C.lib().sfv_config_defaults(cfg, nx, ny, dx, dy)
    cfg.bc = C.SFV_BC_PERIODIC
    cfg.friction = law
    cfg.Cz = Cz
    cfg.stepper = stepper
    cfg.bottom_source = 1
    cfg.desingularizing_kappa = kappa
    cfg.depth_cutoff = cutoff
    return cfg


def _run(cfg, Bval, h, hu, hv, max_dt):
    nxT, nyT = cfg.nx + 2 * cfg.gc, cfg.ny + 2 * cfg.gc
    B = np.full((nxT + 1, nyT + 1), Bval, order="F")
    o = O.Oracle2D(cfg, B)
    U = np.empty((cfg.nx, cfg.ny, 3), order="F")
    U[..., 0], U[..., 1], U[..., 2] = Bval + h, hu, hv
    o.set_interior(U)
    dt = o.step(max_dt, 0.0)
    return dt, o.state()


def _check_uniform(S, w, hu, hv, gc=2):
    # interior cells (the domain corners hold ghosts no boundary condition writes, bc.jl:81-110)
    for v, want in enumerate((w, hu, hv)):
        got = S[gc:-gc, gc:-gc, v]
        assert np.all(np.abs(got - want) <= RTOL * max(abs(want), 1e-300)), (v, got.ravel()[:3], want)


LAWS = [C.SFV_FRICTION_BH2021, C.SFV_FRICTION_BSA2012, C.SFV_FRICTION_FCG2016]


@pytest.mark.parametrize("law", LAWS)
def test_forward_euler_one_implicit_step(law):
    Cz, h, hu, hv, dt = 0.05 ** 2, 0.37, 0.21, -0.13, 0.01
    cfg = _cfg(law, C.SFV_FORWARD_EULER, Cz=Cz)
    got_dt, S = _run(cfg, 0.75, h, hu, hv, dt)
    assert got_dt == dt                       # max_dt binds: dt = min(cfl*..., max_dt)
    f = _factor(law, Cz, h, hu, hv, 1e-5)
    _check_uniform(S, 0.75 + h, hu / (1 - dt * f), hv / (1 - dt * f))


@pytest.mark.parametrize("law", LAWS)
def test_runge_kutta2_friction_uses_the_stage_input(law):
    Cz, h, hu, hv, dt = 0.04 ** 2, 0.29, -0.17, 0.08, 0.02
    cfg = _cfg(law, C.SFV_RUNGE_KUTTA2, Cz=Cz)
    got_dt, S = _run(cfg, 0.75, h, hu, hv, dt)
    assert got_dt == dt
    f0 = _factor(law, Cz, h, hu, hv, 1e-5)
    hu1, hv1 = hu / (1 - dt * f0), hv / (1 - dt * f0)
    f1 = _factor(law, Cz, h, hu1, hv1, 1e-5)          # stage 2 uses U¹, its input
    _check_uniform(S, 0.75 + h, 0.5 * (hu + hu1) / (1 - dt * f1), 0.5 * (hv + hv1) / (1 - dt * f1))
    # the alternatives a plausible slip would give are far outside the tolerance
    f_out = _factor(law, Cz, h, 0.5 * (hu + hu1), 0.5 * (hv + hv1), 1e-5)
    wrong = 0.5 * (hu + hu1) / (1 - dt * f_out)
    assert abs(S[2, 2, 1] - wrong) > 1e3 * RTOL * abs(wrong)


@pytest.mark.parametrize("law", LAWS)
def test_cfl_timestep_and_friction_together(law):
    # max_dt = Inf: dt from the CFL rule on the uniform flow's wave speeds |u| + sqrt(g h)
    # (centralupwind.jl:48-77: a+ = max(u+c, 0), a- = min(u-c, 0), speed max(|a+|, |a-|))
    Cz, h, hu, hv = 0.03 ** 2, 1.7, 0.9, 0.4
    cfg = _cfg(law, C.SFV_RUNGE_KUTTA2, Cz=Cz, dx=2.0, dy=0.5)
    got_dt, S = _run(cfg, 0.0, h, hu, hv, math.inf)
    c = math.sqrt(9.81 * h)
    want_dt = 0.25 * min(2.0 / (abs(hu / h) + c), 0.5 / (abs(hv / h) + c))
    assert abs(got_dt - want_dt) <= 4e-16 * want_dt
    f0 = _factor(law, Cz, h, hu, hv, 1e-5)
    hu1, hv1 = hu / (1 - got_dt * f0), hv / (1 - got_dt * f0)
    f1 = _factor(law, Cz, h, hu1, hv1, 1e-5)
    _check_uniform(S, h, 0.5 * (hu + hu1) / (1 - got_dt * f1), 0.5 * (hv + hv1) / (1 - got_dt * f1))


@pytest.mark.parametrize("law", LAWS)
def test_speed_zero_leaves_momentum_zero(law):
    cfg = _cfg(law, C.SFV_RUNGE_KUTTA2)
    _, S = _run(cfg, 0.75, 0.5, 0.0, 0.0, 0.05)
    assert np.all(S[..., 1] == 0.0) and np.all(S[..., 2] == 0.0)
    assert np.all(S[2:-2, 2:-2, 0] == 1.25)


@pytest.mark.parametrize("law", LAWS)
def test_one_direction_of_flow(law):
    # u = 0, v != 0: the factor depends on |v| alone and hu stays exactly 0
    Cz, h, hv, dt = 0.02 ** 2, 0.8, 0.3, 0.03
    cfg = _cfg(law, C.SFV_FORWARD_EULER, Cz=Cz)
    _, S = _run(cfg, 0.0, h, 0.0, hv, dt)
    f = -Cz * abs(hv / h) / h ** POWER[law]
    _check_uniform(S, h, 0.0, hv / (1 - dt * f))


@pytest.mark.parametrize("law", LAWS)
def test_thin_layer_uses_the_desingularised_depth(law):
    # h < κ: the friction depth is h* = h²/(2κ) + κ/2 (> h), not h
    kappa, Cz, h, hu, hv, dt = 0.3, 0.05 ** 2, 0.1, 0.02, 0.01, 0.01
    cfg = _cfg(law, C.SFV_FORWARD_EULER, Cz=Cz, kappa=kappa)
    _, S = _run(cfg, 0.75, h, hu, hv, dt)
    hs = h * h / (2 * kappa) + kappa / 2
    assert hs > h
    f = -Cz * math.hypot(hu / hs, hv / hs) / hs ** POWER[law]
    _check_uniform(S, 0.75 + h, hu / (1 - dt * f), hv / (1 - dt * f))
    f_raw = -Cz * math.hypot(hu / h, hv / h) / h ** POWER[law]   # the un-desingularised slip
    assert abs(S[2, 2, 1] - hu / (1 - dt * f_raw)) > 1e-6


@pytest.mark.parametrize("law", LAWS)
def test_one_dimensional_friction(law):
    # friction.jl:45-60: speed = sqrt(u^2) = |u|; uniform periodic 1D flow, RK2
    Cz, h, hu, dt = 0.03 ** 2, 0.45, -0.2, 0.02
    cfg = _cfg(law, C.SFV_RUNGE_KUTTA2, Cz=Cz)
    cfg.nx, cfg.ny = 8, 1
    o = O.Oracle1D(cfg, np.zeros(cfg.nx + 2 * cfg.gc + 1))
    U = np.empty((cfg.nx, 2), order="F")
    U[:, 0], U[:, 1] = h, hu
    o.set_interior(U)
    assert o.step(dt, 0.0) == dt
    S = o.state()
    f0 = -Cz * abs(hu / h) / h ** POWER[law]
    hu1 = hu / (1 - dt * f0)
    f1 = -Cz * abs(hu1 / h) / h ** POWER[law]
    want = 0.5 * (hu + hu1) / (1 - dt * f1)
    assert np.all(np.abs(S[2:-2, 1] - want) <= RTOL * abs(want))
    assert np.all(S[2:-2, 0] == h)


def test_default_law_and_coefficient_are_bh2021_with_cz_0_03_squared():
    # friction.jl:26: ImplicitFriction(; Cz=0.03^2, friction_function=friction_bh2021)
    import volumefluxes as VF
    fr = VF.ImplicitFriction()
    assert fr.Cz == 0.03 ** 2
    cfg = VF.build_config(
        VF.ConservedSystem(None, VF.LinearReconstruction(), None,
                           VF.ShallowWaterEquations(VF.ConstantBottomTopography()),
                           VF.CartesianGrid(4, 4), [], fr),
        VF.RungeKutta2(), VF.CartesianGrid(4, 4), 0.25, 0)
    assert cfg.friction == C.SFV_FRICTION_BH2021 and cfg.Cz == 0.03 ** 2
