"""1D shallow-water runs on the GPU (the 2D kernels on the exact (nx, 3)
embedding) against the 1D oracle, bitwise: the reference's 1D tests
(test_swe1D_sim.jl, test_lake_at_rest.jl) run through this path."""
import numpy as np
import pytest

import volumefluxes as VF

from oracle_runs import oracle_1d_for, scenarios_1d, system_1d

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", scenarios_1d(), ids=lambda c: c[0])
def test_1d_steps_bitwise(case):
    name, system, grid, stepper, U0, cfl = case
    o = oracle_1d_for(system, grid, stepper, U0, cfl=cfl)
    sim = VF.Simulator(VF.make_hip_backend(), system, stepper, grid, cfl=cfl)
    VF.set_current_state(sim, U0)
    t = 0.0
    for k in range(40):
        dr = o.step(float("inf"), t)
        dh = VF.perform_step(sim, float("inf"))
        assert dh == dr, (name, k)
        t += dr
        sim.t += dh
    assert np.array_equal(sim.get_full_state(), o.state()), name


def test_stoker_1024_to_20s_bitwise():
    """The 1D workload: 1024 cells on [0, 1000], h = 2 | 1, RK2, CFL 0.25, t = 20 s."""
    g = VF.CartesianGrid(1024, gc=2, boundary=VF.WallBC(), extent=[[0.0, 1000.0]])
    xc = g.cell_centers()
    U0 = np.zeros((1024, 2)); U0[:, 0] = np.where(xc < 500.0, 2.0, 1.0)
    system = system_1d(g)
    o = oracle_1d_for(system, g, VF.RungeKutta2(), U0)
    o.simulate_to_time(20.0)
    sim = VF.Simulator(VF.make_hip_backend(), system, VF.RungeKutta2(), g, cfl=0.25)
    VF.set_current_state(sim, U0)
    VF.simulate_to_time(sim, 20.0)
    assert sim.t == 20.0
    assert np.array_equal(sim.get_full_state(), o.state())
    st = VF.current_interior_state(sim)
    assert st.h.shape == (1024,) and np.all(st.h > 0.9)


def test_1d_members_in_an_ensemble():
    cases = [c for c in scenarios_1d() if c[3].__class__ is VF.RungeKutta2]
    sims, refs = [], []
    for name, system, grid, stepper, U0, cfl in cases:
        sim = VF.Simulator(VF.make_hip_backend(), system, stepper, grid, cfl=cfl)
        VF.set_current_state(sim, U0)
        sims.append(sim)
        o = oracle_1d_for(system, grid, stepper, U0, cfl=cfl)
        t = 0.0
        for _ in range(16):
            t += o.step(float("inf"), t)
        refs.append(o.state())
    VF.Ensemble(sims).advance_steps(16)
    for sim, ref, c in zip(sims, refs, cases):
        assert np.array_equal(sim.get_full_state(), ref), c[0]


def _bump_run(system, grid, stepper, T=0.05, elevate=0.0):
    """test_swe1D_sim.jl:26-41 / test_friction_1d.jl:26-41: a Gaussian bump of
    water, θ = 1.05, CFL 0.2, to T = 0.05 with match_endtime."""
    xc = grid.cell_centers()
    U0 = np.zeros((grid.nx, 2))
    U0[:, 0] = np.exp(-(xc - 0.5) ** 2 / 0.001) + 1.5 + elevate
    sim = VF.Simulator(VF.make_hip_backend(), system, stepper, grid, cfl=0.2)
    VF.set_current_state(sim, U0)
    VF.simulate_to_time(sim, T)
    assert sim.t == T
    return VF.current_interior_state(sim)


@pytest.mark.parametrize("law", ["none", "friction"])
@pytest.mark.parametrize("bottom", ["constant", "field"])
def test_raising_bottom_and_water_by_one_changes_nothing(law, bottom):
    """The practical-equation part of test_swe1D_sim.jl:99-124 and
    test_friction_1d.jl:94-119: B = 1 (constant or a field on the faces) with
    the water raised by 1 gives the B = 0 depths and momenta, Σ|Δ| within the
    reference's 1e-6 (1e-2 with friction and RK2)."""
    grid = VF.CartesianGrid(1024, gc=2)
    stepper = VF.ForwardEulerStepper() if law == "none" else VF.RungeKutta2()
    fr = VF.ImplicitFriction() if law == "friction" else None
    rec = VF.LinearReconstruction(1.05)

    def system(B, sources):
        eq = VF.ShallowWaterEquations1D(B)
        return VF.ConservedSystem(None, rec, VF.CentralUpwind(eq), eq, grid, sources, fr)

    flat = _bump_run(system(VF.ConstantBottomTopography(0.0), []), grid, stepper)
    B = (VF.ConstantBottomTopography(1.0) if bottom == "constant"
         else VF.BottomTopography1D(np.ones(grid.nx + 2 * grid.gc + 1), None, grid))
    raised = _bump_run(system(B, [VF.SourceTermBottom()]), grid, stepper, elevate=1.0)
    tol = 1e-6 if law == "none" else 1e-2
    assert np.sum(np.abs(flat.h - (raised.h - 1.0))) <= tol
    assert np.sum(np.abs(flat.hu - raised.hu)) <= tol
