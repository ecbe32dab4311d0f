"""The 1D runs of the HIP backend are the 2D scheme on an (nx, 3) grid
(volumefluxes.model.embed_1d).  The reference checks that 2D on (n, 3)
equals 1D bitwise, dt included (test_swashes.jl:317-328); this pins the
embedding the backend uses (huge dy, replicated bottom and factor, tables
and friction) on the CPU oracle: the 2D oracle on the embedded system must
equal the 1D oracle bit for bit in the middle row, with equal dt."""
import numpy as np
import pytest

import volumefluxes as VF
from volumefluxes import _capi as C

from oracle_runs import oracle_1d_for, scenarios_1d
import oracle as O


@pytest.mark.parametrize("case", scenarios_1d(), ids=lambda c: c[0])
def test_embedded_2d_equals_1d(case):
    name, system, grid, stepper, U0, cfl = case
    o1 = oracle_1d_for(system, grid, stepper, U0, cfl=cfl)
    g2, s2 = VF.embed_1d(system, grid)
    cfg = VF.build_config(s2, stepper, g2, cfl, C.SFV_F64)
    o2 = O.Oracle2D(cfg, np.asfortranarray(s2.equation.B.corners(g2)))
    o2._keep = cfg
    U2 = np.zeros((grid.nx, 3, 3)); U2[:, :, :2] = U0[:, None, :]
    o2.set_interior(U2)
    t1 = t2 = 0.0
    for k in range(60):
        d1 = o1.step(float("inf"), t1)
        d2 = o2.step(float("inf"), t2)
        assert d1 == d2, (name, k, d1, d2)
        t1 += d1
        t2 += d2
    A, Bm = o1.state(), o2.state()[:, grid.gc + 1, :2]
    assert np.array_equal(A, Bm), name
    assert np.all(np.isfinite(A))


def test_embedding_geometry():
    g = VF.CartesianGrid(100, gc=2, boundary=VF.WallBC(), extent=[[0.0, 10.0]])
    sysd = VF.ConservedSystem(None, VF.LinearReconstruction(), VF.CentralUpwind(VF.ShallowWaterEquations1D()),
                              VF.ShallowWaterEquations1D(), g, [])
    g2, s2 = VF.embed_1d(sysd, g)
    assert (g2.nx, g2.ny, g2.gc) == (100, 3, 2)
    assert g2.delta[0] == g.delta[0] and g2.delta[1] >= g.delta[0] * 2.0 ** 59
