"""Scenario builders shared by the tests: each scenario is a reference-style
system description (grid, equation, sources, stepper) plus an initial
interior state, runnable on the HIP backend and on the CPU oracle."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

import volumefluxes as VF
from volumefluxes import _capi as C
from volumefluxes.scenarios import dam_break_initial_state, terrain_initial_state, terrain_topography


@dataclass
class Scenario:
    name: str
    grid: VF.CartesianGrid
    system: VF.ConservedSystem
    stepper: object
    cfl: float
    U0: np.ndarray                 # interior (nx, ny, 3)
    realtype: type = np.float64
    B: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    @property
    def precision(self):
        return C.SFV_F32_STORAGE if self.realtype == np.float32 else C.SFV_F64

    def config(self):
        return VF.build_config(self.system, self.stepper, self.grid, self.cfl, self.precision)

    def corners(self):
        return np.asfortranarray(self.system.equation.B.corners(self.grid), dtype=np.float64)


def make_oracle(sc: Scenario):
    from oracle import Oracle2D  # test infrastructure
    cfg = sc.config()
    o = Oracle2D(cfg, sc.corners())
    o._cfg_keep = cfg
    o.set_interior(sc.U0)
    return o


def make_hip(sc: Scenario, kernels: str | None = None):
    backend = VF.HIPBackend(sc.realtype)
    sim = VF.Simulator(backend, sc.system, sc.stepper, sc.grid, cfl=sc.cfl)
    if kernels is not None:
        sim.set_kernels(kernels)
    VF.set_current_state(sim, sc.U0)
    return sim


def _system(grid, B, *, sources, friction=None, rec=None, kappa=1e-5, cutoff=1e-5):
    topo = VF.BottomTopography2D(B, None, grid) if B is not None else VF.ConstantBottomTopography()
    eq = VF.ShallowWaterEquations(topo, depth_cutoff=cutoff, desingularizing_kappa=kappa)
    return VF.ConservedSystem(None, rec if rec is not None else VF.LinearReconstruction(), VF.CentralUpwind(eq), eq,
                              grid, sources, friction)


def terrain(nx, ny, *, seed=1, bc=None, stepper=None, realtype=np.float64, gc=2, dx=1.0, w0=2.0, rain=0.036,
            friction=True, kappa=1e-4, name=None, n_bumps=None, rec=None, width=None):
    grid = VF.CartesianGrid(nx, ny, gc=gc, boundary=bc or VF.WallBC(), extent=[[0.0, nx * dx], [0.0, ny * dx]])
    nb = n_bumps if n_bumps is not None else 32
    if width is None:   # bump widths scaled to the domain, so small grids still see terrain
        lo = max(2.0, min(nx, ny) * dx / 40)
        width = (lo, max(lo, max(nx, ny) * dx / 8))
    B = terrain_topography(grid, seed=seed, n_bumps=nb, amp=(0.5, 5.0), width=width)
    sources = [VF.SourceTermBottom()]
    if rain:
        sources.append(VF.ConstantRain(rain))
    fr = VF.ImplicitFriction(Cz=0.03 * 0.03) if friction else None
    system = _system(grid, B, sources=sources, friction=fr, kappa=kappa, rec=rec)
    return Scenario(name or f"terrain{nx}x{ny}", grid, system, stepper or VF.RungeKutta2(), 0.25,
                    terrain_initial_state(grid, B, w0), realtype, B)


def dam_break(nx, ny, *, hl=0.005, hr=0.001, L=10.0, stepper=None, bc=None, direction=0, cfl=0.1,
              kappa=1e-7, cutoff=1e-7, name=None):
    ext = [[0.0, L], [0.0, 5.0]] if direction == 0 else [[0.0, 5.0], [0.0, L]]
    grid = VF.CartesianGrid(nx, ny, gc=2, boundary=bc or VF.WallBC(), extent=ext)
    system = _system(grid, None, sources=[], kappa=kappa, cutoff=cutoff)
    U0 = dam_break_initial_state(grid, L / 2, hl, hr, direction)
    return Scenario(name or f"dambreak{nx}x{ny}", grid, system, stepper or VF.RungeKutta2(), cfl, U0)


def run_both_compare(sc: Scenario, nsteps: int, *, kernels=None, max_dt=float("inf")):
    """Step HIP and oracle nsteps each; return the list of (dt_hip, dt_ref)
    and the final full states."""
    o = make_oracle(sc)
    sim = make_hip(sc, kernels)
    dts = []
    t = 0.0
    for _ in range(nsteps):
        dr = o.step(max_dt, t)
        dh = VF.perform_step(sim, max_dt)
        sim.t = float(sc.realtype(sc.realtype(sim.t) + sc.realtype(dh)))
        t = float(sc.realtype(sc.realtype(t) + sc.realtype(dr)))
        dts.append((dh, dr))
    return dts, sim.get_full_state(), o.state()


def max_rel_diff(a: np.ndarray, b: np.ndarray) -> float:
    """max|Δ| / max|ref| over each variable plane, the worst plane."""
    worst = 0.0
    for v in range(a.shape[-1]):
        ref = np.max(np.abs(b[..., v]))
        d = np.max(np.abs(a[..., v].astype(np.float64) - b[..., v].astype(np.float64)))
        worst = max(worst, d / ref if ref > 0 else d)
    return worst
