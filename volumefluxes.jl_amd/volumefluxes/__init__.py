"""volumefluxes — MI355X (gfx950) backend for the 2D shallow-water path of
sintefmath/VolumeFluxes.jl (SinFVM), with the reference's Simulator /
backend / simulate_to_time API mirrored in Python over the C ABI of
libsinfvm_hip.so (include/sinfvm_hip.h).

    import volumefluxes as VF
    backend = VF.make_hip_backend()
    grid = VF.CartesianGrid(nx, ny, gc=2, boundary=VF.WallBC(), extent=[[0, Lx], [0, Ly]])
    eq = VF.ShallowWaterEquations(VF.BottomTopography2D(B, backend, grid))
    system = VF.ConservedSystem(backend, VF.LinearReconstruction(), VF.CentralUpwind(eq), eq, grid,
                                [VF.SourceTermBottom(), VF.ConstantRain(0.036)], VF.ImplicitFriction())
    sim = VF.Simulator(backend, system, VF.RungeKutta2(), grid, cfl=0.25)
    VF.set_current_state(sim, initial)          # (nx, ny, 3): w, hu, hv
    VF.simulate_to_time(sim, 600.0)
"""
from ._capi import (ArgumentError, BudgetError, DomainError, NonFiniteTimestepError, SfvAssertionError,
                    SfvError)
from .grid import (XDIR, YDIR, CartesianGrid, NeumannBC, PeriodicBC, WallBC, cell_centers, cell_faces,
                   compute_dx)
from .model import (BottomTopography1D, BottomTopography2D, CentralUpwind, ConservedSystem, ConstantBottomTopography,
                    ConstantInfiltration, ConstantRain, ForwardEulerStepper, FunctionalRain,
                    HortonInfiltration, ImplicitFriction, LinearReconstruction, NoReconstruction, RungeKutta2,
                    ShallowWaterEquations, ShallowWaterEquations1D, SourceTermBottom, TimeDependentRain, build_config, embed_1d, rain_field_values,
                    collect_topography_cells, compute_infiltration, compute_rain, friction_bh2021,
                    friction_bsa2012, friction_fcg2016, number_of_substeps, validate_bottom_topography)
from .simulator import (HIPBackend, IntervalWriter, MultipleCallbacks, Simulator, advance_steps,
                        current_interior_state, current_state, current_time, current_timestep, device_count,
                        get_available_backends, has_hip_backend, make_hip_backend, name, perform_step,
                        set_current_state, simulate_to_time)

from . import io, model  # noqa: E402
from .io import SnapshotWriter, TotalWaterVolume, bottom_per_cell, coarsen, load_grid  # noqa: E402
from .adjoint import adjoint_gradient, manning_gradient  # noqa: E402
from .batch import Ensemble  # noqa: E402
from .serving import Server  # noqa: E402

__all__ = [n for n in dir() if not n.startswith("_")]
