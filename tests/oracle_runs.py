"""Helpers that run reference-style scenarios on the CPU oracle (test
infrastructure; no GPU)."""
from __future__ import annotations

import math

import numpy as np

import oracle as O
import volumefluxes as VF
from volumefluxes import _capi as C


def system_2d(grid, B=None, *, sources=(), friction=None, rec=None, rho=1.0, g=9.81, cutoff=1e-5, kappa=1e-5):
    topo = VF.BottomTopography2D(B, None, grid) if B is not None else VF.ConstantBottomTopography()
    eq = VF.ShallowWaterEquations(topo, rho=rho, g=g, depth_cutoff=cutoff, desingularizing_kappa=kappa)
    rec = rec if rec is not None else VF.LinearReconstruction()
    return VF.ConservedSystem(None, rec, VF.CentralUpwind(eq), eq, grid, list(sources), friction)


def oracle_2d(grid, system, stepper, U0, *, cfl=0.25, t0=0.0, realtype=np.float64):
    prec = C.SFV_F32_STORAGE if realtype == np.float32 else C.SFV_F64
    cfg = VF.build_config(system, stepper, grid, cfl, prec)
    o = O.Oracle2D(cfg, np.asfortranarray(system.equation.B.corners(grid)))
    o._keep = cfg
    o.set_interior(U0)
    o.t = t0
    return o


def config_1d(nx, L, *, x0=0.0, gc=2, bc=C.SFV_BC_WALL, theta=1.2, stepper=C.SFV_RUNGE_KUTTA2, cfl=0.25,
              cutoff=1e-5, kappa=1e-5, bottom=0, rec=C.SFV_LINEAR_RECONSTRUCTION):
    cfg = C.SfvConfig()
    cfg.nx, cfg.ny, cfg.gc = nx, 1, gc
    cfg.dx = cfg.dy = (L - x0) / nx
    cfg.bc = bc
    cfg.g, cfg.rho, cfg.depth_cutoff, cfg.desingularizing_kappa = 9.81, 1.0, cutoff, kappa
    cfg.reconstruction, cfg.theta = rec, theta
    cfg.stepper, cfg.cfl = stepper, cfl
    cfg.bottom_source = bottom
    cfg.rain_before_infiltration = 1
    return cfg


def faces_1d(nx, x0, x1, gc, interior=True):
    dx = (x1 - x0) / nx
    if interior:
        return VF.grid._linrange(x0, x1, nx + 1)
    return VF.grid._linrange(x0 - gc * dx, x1 + gc * dx, nx + 2 * gc + 1)


def centres_1d(nx, x0, x1):
    f = faces_1d(nx, x0, x1, 0)
    return f[:-1] + (f[1] - f[0]) / 2.0


def stoker_cm(hl, hr, g=9.81):
    """Intermediate celerity cm = sqrt(g hm) of the wet dam break (Stoker;
    SWASHES 4.1.1): the depth hm where the left rarefaction's velocity
    2(sqrt(g hl) - sqrt(g hm)) equals the velocity behind the right shock
    (hm - hr) sqrt(g (hm + hr) / (2 hm hr)); unique in (hr, hl), by bisection."""
    cl = math.sqrt(g * hl)
    f = lambda hm: 2 * (cl - math.sqrt(g * hm)) - (hm - hr) * math.sqrt(g * (hm + hr) / (2 * hm * hr))
    lo, hi = hr, hl
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            lo = mid
        else:
            hi = mid
    return math.sqrt(g * 0.5 * (lo + hi))


def dam_break_exact(x, t, x0, hl, hr, g=9.81):
    """Stoker (hr > 0) / Ritter (hr = 0) dam break: depth and velocity."""
    x = np.asarray(x, dtype=float)
    cl = math.sqrt(g * hl)
    xA = x0 - t * cl
    h = np.full_like(x, hr)
    u = np.zeros_like(x)
    if hr > 0:
        cm = stoker_cm(hl, hr, g)
        xB = x0 + t * (2 * cl - 3 * cm)
        xC = x0 + t * (2 * cm * cm * (cl - cm)) / (cm * cm - g * hr)
    else:
        cm = 0.0
        xB = x0 + 2 * t * cl
        xC = xB
    rare = (x >= xA) & (x < xB)
    h[x < xA] = hl
    h[rare] = (4.0 / (9.0 * g)) * (cl - (x[rare] - x0) / (2 * t)) ** 2
    u[rare] = (2.0 / 3.0) * ((x[rare] - x0) / t + cl)
    mid = (x >= xB) & (x < xC)
    h[mid] = cm * cm / g
    u[mid] = 2 * (cl - cm)
    return h, u


def oracle_1d_for(system, grid, stepper, U0, *, cfl=0.25):
    """The 1D oracle (the reference's CartesianGrid{1} scheme) for a 1D
    system description: the configuration of the 2D embedding with the grid
    put back to 1D (ny = 1) and the 1D bottom and factor field."""
    g2, s2 = VF.embed_1d(system, grid)
    cfg = VF.build_config(s2, stepper, g2, cfl, C.SFV_F64)
    cfg.ny = 1
    cfg.dy = cfg.dx
    keep = [cfg]
    fac = [s for s in system.source_terms if isinstance(s, VF.HortonInfiltration)]
    if fac:
        f1 = np.ascontiguousarray(fac[0].factor.reshape(-1), dtype=np.float64)
        cfg.infiltration_factor = C.dptr(f1)
        keep.append(f1)
    o = O.Oracle1D(cfg, np.ascontiguousarray(system.equation.B.corners(grid), dtype=np.float64))
    o._keep = keep
    o.set_interior(U0)
    return o


def system_1d(grid, B=None, *, sources=(), friction=None, rec=None, cutoff=1e-5, kappa=1e-5):
    topo = VF.BottomTopography1D(B, None, grid) if B is not None else VF.ConstantBottomTopography()
    eq = VF.ShallowWaterEquations1D(topo, depth_cutoff=cutoff, desingularizing_kappa=kappa)
    return VF.ConservedSystem(None, rec if rec is not None else VF.LinearReconstruction(), VF.CentralUpwind(eq), eq,
                              grid, list(sources), friction)


def scenarios_1d():
    """1D cases for the 2D-embedding checks: (name, system, grid, stepper, U0, cfl)."""
    out = []
    g = VF.CartesianGrid(256, gc=2, boundary=VF.WallBC(), extent=[[0.0, 1000.0]])
    xc = g.cell_centers()
    U = np.zeros((256, 2)); U[:, 0] = np.where(xc < 500.0, 2.0, 1.0)
    out.append(("stoker", system_1d(g), g, VF.RungeKutta2(), U, 0.25))
    U = np.zeros((256, 2)); U[:, 0] = np.where(xc < 500.0, 2.0, 0.0)
    out.append(("ritter_fe", system_1d(g, cutoff=1e-7, kappa=1e-7), g, VF.ForwardEulerStepper(), U, 0.2))
    gb = VF.CartesianGrid(300, gc=2, boundary=VF.WallBC(), extent=[[0.0, 300.0]])
    xf = gb.cell_faces(interior=False)
    B = 1e-3 * xf + 1.5 * np.exp(-((xf - 150.0) ** 2) / (2 * 20.0 ** 2))
    Bc = 0.5 * (B[:-1] + B[1:])[2:-2]
    U = np.zeros((300, 2)); U[:, 0] = np.maximum(Bc, 1.0)
    out.append(("bump_rain_friction", system_1d(gb, B, sources=[VF.SourceTermBottom(), VF.ConstantRain(36.0)],
                                                friction=VF.ImplicitFriction(Cz=0.03 ** 2), kappa=1e-4),
                gb, VF.RungeKutta2(), U, 0.25))
    gp = VF.CartesianGrid(200, gc=2, boundary=VF.PeriodicBC(), extent=[[0.0, 100.0]])
    xcp = gp.cell_centers()
    fac = np.linspace(0.5, 1.5, 204)
    U = np.zeros((200, 2)); U[:, 0] = 1.0 + 0.3 * np.sin(2 * np.pi * xcp / 100.0); U[:, 1] = 0.2
    out.append(("periodic_horton_table", system_1d(gp, sources=[VF.TimeDependentRain([10.0, 50.0, 0.0], [0.0, 2.0, 4.0]),
                                                                 VF.HortonInfiltration(gp, factor=fac)]),
                gp, VF.RungeKutta2(), U, 0.25))
    return out
