"""The forward-mode restatement in the oracle (oracle/oracle_ad.hpp), used
to check the GPU adjoint: on plain doubles it must equal the oracle's step
bit for bit; its directional derivative of J must match central finite
differences of the oracle."""
import numpy as np
import pytest

import oracle as O
import volumefluxes as VF
from helpers import dam_break, make_oracle, terrain


def _cfg(sc):
    cfg = sc.config()
    cfg._keep = cfg
    return cfg


def _oracle_run(sc, nsteps, B=None, Cz=None):
    cfg = _cfg(sc)
    if Cz is not None:
        cfg.Cz = Cz
    o = O.Oracle2D(cfg, sc.corners() if B is None else np.asfortranarray(B))
    o._k = cfg
    o.set_interior(sc.U0)
    t, dts = 0.0, []
    for _ in range(nsteps):
        dt = o.step(float("inf"), t)
        dts.append(dt)
        t += dt
    return o.state(), np.array(dts)


CASES = [lambda: terrain(36, 28, seed=5), lambda: terrain(30, 22, seed=6, bc=VF.PeriodicBC()),
         lambda: dam_break(40, 9, hr=0.0)]


@pytest.mark.parametrize("mk", CASES)
def test_ad_forward_equals_oracle_bitwise(mk):
    sc = mk()
    U_ref, dts_ref = _oracle_run(sc, 25)
    U, dts = O.ad_forward(_cfg(sc), sc.corners(), sc.U0, 25)
    assert np.array_equal(dts, dts_ref)
    assert np.array_equal(U, U_ref)


def _J(U, B, target, gc):
    from volumefluxes.scenarios import cell_bottom
    h = U[gc:-gc, gc:-gc, 0] - cell_bottom(B)[gc:-gc, gc:-gc]
    return float(np.sum((h - target) ** 2))


def smooth_flow(sc):
    """An all-wet state with no exact ties in the scheme's branches (no zero
    velocities, no flat water): every derivative is a classical one."""
    from volumefluxes.scenarios import cell_bottom
    Bc = cell_bottom(sc.corners())[sc.grid.gc:-sc.grid.gc, sc.grid.gc:-sc.grid.gc]
    X, Y = np.meshgrid(np.arange(sc.grid.nx), np.arange(sc.grid.ny), indexing="ij")
    U = np.zeros((sc.grid.nx, sc.grid.ny, 3))
    U[:, :, 0] = Bc.max() + 1.0 + 0.2 * np.sin(X / 3.0 + 0.4) * np.cos(Y / 4.0 + 0.2)
    U[:, :, 1] = 0.3 * np.cos(X / 5.0 + Y / 7.0 + 0.3) + 0.05
    U[:, :, 2] = 0.2 * np.sin(X / 6.0 - Y / 3.0 + 0.7) + 0.03
    sc.U0 = U
    return sc


def _fd(sc, nsteps, target, eps, dB=None, dU=None, dCz=None):
    B = sc.corners()
    Cz = sc.config().Cz
    out = []
    for sgn in (1, -1):
        Bp = B + sgn * eps * dB if dB is not None else B
        s0 = sc.U0
        if dU is not None:
            sc.U0 = s0 + sgn * eps * dU
        U, _ = _oracle_run(sc, nsteps, B=Bp, Cz=Cz + sgn * eps * dCz if dCz is not None else None)
        sc.U0 = s0
        out.append(_J(U, Bp, target, sc.grid.gc))
    return (out[0] - out[1]) / (2 * eps)


def test_ad_directional_derivative_matches_finite_differences():
    sc = smooth_flow(terrain(32, 24, seed=9))
    nsteps = 12
    B = sc.corners()
    from volumefluxes.scenarios import cell_bottom
    target = sc.U0[:, :, 0] - cell_bottom(B)[2:-2, 2:-2] - 0.3
    rng = np.random.default_rng(3)
    dB = rng.normal(size=B.shape) * 1e-2
    dU = rng.normal(size=sc.U0.shape) * 1e-2
    J, dJ, U, _ = O.ad_run(_cfg(sc), B, sc.U0, nsteps, target, dB=dB)
    assert abs(J - _J(U, B, target, 2)) <= 1e-12 * abs(J)
    fd = _fd(sc, nsteps, target, 1e-5, dB=dB)
    assert abs(dJ - fd) <= 1e-6 * abs(fd)
    _, dJu, _, _ = O.ad_run(_cfg(sc), B, sc.U0, nsteps, target, dU0=dU)
    fdu = _fd(sc, nsteps, target, 1e-5, dU=dU)
    assert abs(dJu - fdu) <= 1e-5 * abs(fdu)
    Cz = sc.config().Cz
    _, dJc, _, _ = O.ad_run(_cfg(sc), B, sc.U0, nsteps, target, dCz=1.0)
    # Cz moves J by ~1e-8 over these steps while rounding changes of the whole
    # trajectory move it by ~1e-12: central differences resolve ~1e-4 here
    fdc = _fd(sc, nsteps, target, 5e-2 * Cz, dCz=1.0)
    assert abs(dJc - fdc) <= 2e-4 * abs(fdc)
