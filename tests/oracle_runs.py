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
