# HIPBackend.jl — VolumeFluxes backend that runs the shallow-water step on an
# AMD MI355X (gfx950) through libsinfvm_hip.so (include/sinfvm_hip.h).
#
# Load after VolumeFluxes (the reference's module, src/SinFVM.jl:21):
#     using VolumeFluxes
#     include("HIPBackend.jl"); using .VolumeFluxesHIP
#     backend = VolumeFluxesHIP.make_hip_backend()      # or make_hip_backend(Float32)
#     sim = Simulator(backend, system, RungeKutta2(), grid)
#     set_current_state!(sim, initial); simulate_to_time(sim, T)
#
# How it plugs in (src/backends/kernel_abstractions.jl:24-55, src/simulator.jl):
#   * HIPBackend{T} <: VolumeFluxes.Backend with field `realtype`, like
#     KernelAbstractionBackend.  The reference's generic create_buffer /
#     convert_to_backend (src/backends/buffer.jl:22,29-31) give it plain host
#     arrays: every Volume of a HIP simulator is a host mirror.
#   * perform_step!(sim::Simulator{<:HIPBackend}, max_dt) is one sfv_step call
#     (src/simulator.jl:145-177).  simulate_to_time without a callback is one
#     sfv_run call (a device-resident loop); with a callback the reference's
#     own loop runs and calls perform_step! per step.
#   * set_current_state!, current_state, current_interior_state sync the host
#     mirror (src/simulator.jl:98-135).
#   * A CartesianGrid{1} system (ShallowWaterEquations1D, src/equations/swe_1D.jl)
#     runs as the library's exact (nx, 3) embedding (sfv_config.embedded_1d):
#     every row carries the 1D state, dy = dx * 2^60, and the middle row is the
#     1D scheme bit for bit (the reference's own check, test/test_swashes.jl:317-328).
# The device handle of a simulator is created on first use and found again by
# the IDENTITY of its `t` MVector (the Simulator struct itself is immutable, and
# `t` is mutated every step, so a content-keyed table would lose the handle).
# The handle is destroyed when that MVector is garbage-collected.
module VolumeFluxesHIP

using VolumeFluxes
using StaticArrays

const libsinfvm = get(ENV, "SINFVM_HIP_LIB", joinpath(@__DIR__, "..", "lib", "libsinfvm_hip.so"))

# ---- status codes and enums (include/sinfvm_hip.h) ------------------------
const SFV_OK = 0
const SFV_ERR_DOMAIN = 1
const SFV_ERR_ARGUMENT = 2
const SFV_ERR_ASSERTION = 3
const SFV_ERR_NONFINITE_DT = 5
const SFV_BC_WALL, SFV_BC_PERIODIC, SFV_BC_NEUMANN = 0, 1, 2
const SFV_FORWARD_EULER, SFV_RUNGE_KUTTA2 = 1, 2
const SFV_LAYOUT_FULL, SFV_LAYOUT_INTERIOR = 0, 1

# Field order and types match `sfv_config` exactly (C layout of an isbits struct).
mutable struct SfvConfig
    nx::Int32; ny::Int32; gc::Int32
    dx::Float64; dy::Float64
    bc::Int32
    g::Float64; rho::Float64; depth_cutoff::Float64; desingularizing_kappa::Float64
    reconstruction::Int32
    theta::Float64
    stepper::Int32
    cfl::Float64
    bottom_source::Int32
    friction::Int32
    Cz::Float64
    rain::Int32
    rain_rate::Float64
    rain_table_len::Int32
    rain_table_rates::Ptr{Float64}
    rain_table_times::Ptr{Float64}
    infiltration::Int32
    infiltration_rate::Float64
    horton_f0::Float64; horton_fc::Float64; horton_k::Float64
    infiltration_factor_scalar::Float64
    infiltration_factor::Ptr{Float64}
    rain_before_infiltration::Int32
    precision::Int32
    embedded_1d::Int32
    SfvConfig() = new()
end

struct HIPBackend{RealType} <: VolumeFluxes.Backend
    realtype::Type{RealType}
end

device_count() = Int(ccall((:sfv_device_count, libsinfvm), Cint, ()))

function make_hip_backend(RealType=Float64)
    RealType in (Float64, Float32) || throw(ArgumentError("HIPBackend supports Float64 and Float32"))
    device_count() > 0 || error("no HIP device visible to libsinfvm_hip")
    return HIPBackend{RealType}(RealType)
end
VolumeFluxes.name(::HIPBackend) = "HIP"

"get_available_backends() plus the HIP backend when a device is present (kernel_abstractions.jl:45-55)."
function available_backends_with_hip()
    backends = VolumeFluxes.get_available_backends()
    try
        push!(backends, make_hip_backend())
    catch err
        @show err
    end
    return backends
end

# ---- errors ----------------------------------------------------------------
last_error(h) = unsafe_string(ccall((:sfv_last_error, libsinfvm), Cstring, (Ptr{Cvoid},), h))
create_error() = unsafe_string(ccall((:sfv_create_error, libsinfvm), Cstring, ()))

function check(h, rc)
    rc == SFV_OK && return nothing
    msg = h == C_NULL ? create_error() : last_error(h)
    rc == SFV_ERR_DOMAIN && throw(DomainError(msg))
    rc == SFV_ERR_ARGUMENT && throw(ArgumentError(msg))
    rc == SFV_ERR_ASSERTION && throw(AssertionError(msg))
    rc == SFV_ERR_NONFINITE_DT && (@warn msg; return nothing)
    error("libsinfvm_hip: $msg (status $rc)")
end

# ---- system -> sfv_config ----------------------------------------------------
friction_code(f) = f === VolumeFluxes.friction_bh2021 ? 1 : f === VolumeFluxes.friction_bsa2012 ? 2 :
                   f === VolumeFluxes.friction_fcg2016 ? 3 : throw(ArgumentError("unsupported friction function $f"))
bc_code(::VolumeFluxes.WallBC) = SFV_BC_WALL
bc_code(::VolumeFluxes.PeriodicBC) = SFV_BC_PERIODIC
bc_code(::VolumeFluxes.NeumannBC) = SFV_BC_NEUMANN
stepper_code(::VolumeFluxes.RungeKutta2) = SFV_RUNGE_KUTTA2
stepper_code(::VolumeFluxes.ForwardEulerStepper) = SFV_FORWARD_EULER

mutable struct HandleBox
    h::Ptr{Cvoid}
    host_valid::Bool           # substep_outputs[1] mirrors the device state
end

# objectid(sim.t) => (weak reference to sim.t, its handle).  Identity, not
# content: objectid of a mutable object is its address, stable while it lives.
# The weak reference guards against an address reused after a collection.
const HANDLES = Dict{UInt,Tuple{WeakRef,HandleBox}}()
const HANDLES_LOCK = ReentrantLock()

# sim.t's finalizer: destroy the device handle (a plain ccall; no task switch,
# no Dict mutation inside a finalizer).  The stale entry is swept later.
function release!(box::HandleBox)
    h = box.h
    box.h = C_NULL
    h == C_NULL || ccall((:sfv_destroy, libsinfvm), Cvoid, (Ptr{Cvoid},), h)
    return nothing
end

function sweep_handles!()
    for (k, (wr, box)) in collect(HANDLES)
        (wr.value === nothing || box.h == C_NULL) && delete!(HANDLES, k)
    end
end

function lookup_handle(tvec)
    e = get(HANDLES, objectid(tvec), nothing)
    e === nothing && return nothing
    wr, box = e
    (wr.value === tvec && box.h != C_NULL) ? box : nothing
end

is_1d(sim) = VolumeFluxes.dimension(sim.grid) == 1
const EMBED_ROWS = 3           # rows of the (nx, 3) embedding of a 1D grid

function build_config(sim)
    sys, grid, eq = sim.system, sim.grid, sim.system.equation
    one_d = is_1d(sim)
    sys.numericalflux isa VolumeFluxes.CentralUpwind || throw(ArgumentError("HIPBackend runs CentralUpwind"))
    if one_d
        eq isa VolumeFluxes.ShallowWaterEquations1D ||
            throw(ArgumentError("HIPBackend runs ShallowWaterEquations1D on a CartesianGrid{1}"))
    else
        eq isa VolumeFluxes.ShallowWaterEquations ||
            throw(ArgumentError("HIPBackend runs ShallowWaterEquations on a CartesianGrid{2}"))
    end
    gc = grid.ghostcells[1]
    nx = grid.totalcells[1] - 2gc
    dx = VolumeFluxes.compute_dx(grid, VolumeFluxes.XDIR)
    # 1D: the (nx, 3) embedding; dy = dx * 2^60 keeps the y speeds out of dt
    ny = one_d ? EMBED_ROWS : grid.totalcells[2] - 2gc
    dy = one_d ? dx * 2.0^60 : VolumeFluxes.compute_dx(grid, VolumeFluxes.YDIR)
    c = SfvConfig()
    ccall((:sfv_config_defaults, libsinfvm), Cvoid, (Ref{SfvConfig}, Int32, Int32, Float64, Float64), c,
          nx, ny, dx, dy)
    c.gc = gc
    c.bc = bc_code(grid.boundary)
    c.g, c.rho = eq.g, eq.ρ
    c.depth_cutoff, c.desingularizing_kappa = eq.depth_cutoff, eq.desingularizing_kappa
    if sys.reconstruction isa VolumeFluxes.LinearReconstruction
        c.reconstruction, c.theta = 1, sys.reconstruction.theta
    else
        c.reconstruction = 0
    end
    c.stepper = stepper_code(sim.timestepper)
    c.cfl = sim.cfl
    keep = Any[]
    seen_rain = false
    c.rain_before_infiltration = 1
    nyT = ny + 2gc
    for st in sys.source_terms
        if st isa VolumeFluxes.SourceTermBottom
            c.bottom_source = 1
        elseif st isa VolumeFluxes.ConstantRain
            c.rain, c.rain_rate, seen_rain = 1, st.rain_rate, true
        elseif st isa VolumeFluxes.TimeDependentRain
            rates, times = collect(Float64, st.rain_rates), collect(Float64, st.time)
            push!(keep, rates, times)
            c.rain, c.rain_table_len = 2, length(rates)
            c.rain_table_rates, c.rain_table_times = pointer(rates), pointer(times)
            seen_rain = true
        elseif st isa VolumeFluxes.HortonInfiltration
            f = collect(Float64, st.factor)
            # 1D: the (nxT,) factor repeated on every row of the embedding
            factor = one_d ? repeat(reshape(f, :, 1), 1, nyT) : f
            push!(keep, factor)
            c.infiltration = 2
            c.horton_f0, c.horton_fc, c.horton_k = st.f0, st.fc, st.k
            c.infiltration_factor = pointer(factor)
            seen_rain || (c.rain_before_infiltration = 0)
        elseif st isa VolumeFluxes.ConstantInfiltration
            c.infiltration, c.infiltration_rate = 1, st.infiltration_rate
            seen_rain || (c.rain_before_infiltration = 0)
        else
            throw(ArgumentError("HIPBackend does not run source term $(typeof(st))"))
        end
    end
    fr = sys.implicit_source_term
    if fr !== nothing
        fr isa VolumeFluxes.ImplicitFriction || throw(ArgumentError("unsupported implicit source term"))
        c.friction, c.Cz = friction_code(fr.friction_function), fr.Cz
    end
    c.precision = sim.backend.realtype === Float32 ? 1 : 0
    # B_cell of the embedding takes the 1D lookup 0.5*(B[i]+B[i+1]) (bottom_topography.jl:92)
    c.embedded_1d = one_d ? 1 : 0
    return c, keep
end

# Corner topography of the device grid: (nxT+1, nyT+1); in 1D the reference's
# (nxT+1,) array repeated on every corner row of the embedding.
function corner_topography(sim)
    B = sim.system.equation.B
    nxT = sim.grid.totalcells[1]
    nyT = is_1d(sim) ? EMBED_ROWS + 2sim.grid.ghostcells[1] : sim.grid.totalcells[2]
    B isa VolumeFluxes.ConstantBottomTopography && return fill(Float64(B.B), nxT + 1, nyT + 1)
    Bh = collect(Float64, B.B)
    return is_1d(sim) ? repeat(reshape(Bh, :, 1), 1, nyT + 1) : Bh
end

function handle!(sim::VolumeFluxes.Simulator{<:HIPBackend})
    lock(HANDLES_LOCK) do
        box = lookup_handle(sim.t)
        box === nothing || return box
        sweep_handles!()
        c, keep = build_config(sim)
        href = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve keep check(C_NULL, ccall((:sfv_create, libsinfvm), Cint, (Ref{SfvConfig}, Ref{Ptr{Cvoid}}), c, href))
        box = HandleBox(href[], false)
        HANDLES[objectid(sim.t)] = (WeakRef(sim.t), box)
        finalizer(_ -> release!(box), sim.t)
        B = corner_topography(sim)
        check(box.h, ccall((:sfv_set_topography, libsinfvm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64, Int64),
                           box.h, B, size(B, 1), size(B, 2)))
        # the state the Simulator was built with (zeros, or whatever its mirror holds)
        upload_full!(sim, box)
        return box
    end
end

# The device grid's full state from the host mirror: in 1D every row of the
# embedding carries the (nxT, 2) state, with hv = 0.
function device_full(sim)
    data = sim.substep_outputs[1]._data
    is_1d(sim) || return data
    nyT = EMBED_ROWS + 2sim.grid.ghostcells[1]
    U = zeros(eltype(data), size(data, 1), nyT, 3)
    for j in 1:nyT, v in 1:2
        U[:, j, v] .= view(data, :, v)
    end
    return U
end

function upload_full!(sim, box)
    U = device_full(sim)
    GC.@preserve U check(box.h, ccall((:sfv_set_state, libsinfvm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int32),
                                      box.h, U, SFV_LAYOUT_FULL))
    box.host_valid = true
end

function sync_host!(sim)
    box = handle!(sim)
    box.host_valid && return
    data = sim.substep_outputs[1]._data
    if is_1d(sim)
        # the middle row of the embedding is the 1D state
        nyT = EMBED_ROWS + 2sim.grid.ghostcells[1]
        U = zeros(eltype(data), size(data, 1), nyT, 3)
        check(box.h, ccall((:sfv_get_state, libsinfvm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int32),
                           box.h, U, SFV_LAYOUT_FULL))
        mid = sim.grid.ghostcells[1] + 2
        data[:, 1] .= view(U, :, mid, 1)
        data[:, 2] .= view(U, :, mid, 2)
    else
        check(box.h, ccall((:sfv_get_state, libsinfvm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int32),
                           box.h, data, SFV_LAYOUT_FULL))
    end
    box.host_valid = true
end

# ---- the reference API on a HIP simulator ---------------------------------------
function VolumeFluxes.perform_step!(sim::VolumeFluxes.Simulator{<:HIPBackend}, max_dt)
    box = handle!(sim)
    dt = Ref{Float64}(0.0)
    check(box.h, ccall((:sfv_step, libsinfvm), Cint, (Ptr{Cvoid}, Float64, Float64, Ref{Float64}),
                       box.h, max_dt, sim.t[1], dt))
    sim.current_timestep[1] = dt[]
    box.host_valid = false
    return nothing
end

function VolumeFluxes.simulate_to_time(sim::VolumeFluxes.Simulator{<:HIPBackend}, endtime; match_endtime=true,
                                       callback=nothing, show_progress=true, maximum_timestep=nothing)
    if callback !== nothing
        # the reference's own loop: perform_step! per step, t[1] += dt, callback after each
        return invoke(VolumeFluxes.simulate_to_time, Tuple{VolumeFluxes.Simulator,Any}, sim, endtime;
                      match_endtime=match_endtime, callback=callback, show_progress=show_progress,
                      maximum_timestep=maximum_timestep)
    end
    box = handle!(sim)
    t = Ref{Float64}(sim.t[1])
    steps = Ref{Int64}(0)
    mts = maximum_timestep === nothing ? 0.0 : Float64(maximum_timestep)
    rc = ccall((:sfv_run, libsinfvm), Cint, (Ptr{Cvoid}, Float64, Int32, Float64, Int64, Ref{Float64}, Ref{Int64}),
               box.h, endtime, match_endtime ? 1 : 0, mts, 0, t, steps)
    sim.t[1] = t[]
    if steps[] > 0
        dt = Ref{Float64}(0.0)
        check(box.h, ccall((:sfv_current_timestep, libsinfvm), Cint, (Ptr{Cvoid}, Ref{Float64}), box.h, dt))
        sim.current_timestep[1] = dt[]
    end
    box.host_valid = false
    check(box.h, rc)
    return nothing
end

function VolumeFluxes.current_state(sim::VolumeFluxes.Simulator{<:HIPBackend})
    sync_host!(sim)
    return sim.substep_outputs[1]
end

VolumeFluxes.current_interior_state(sim::VolumeFluxes.Simulator{<:HIPBackend}) =
    VolumeFluxes.interior(VolumeFluxes.current_state(sim))

function VolumeFluxes.set_current_state!(sim::VolumeFluxes.Simulator{<:HIPBackend}, new_state)
    box = handle!(sim)
    sync_host!(sim)
    # the interior written into the mirror as the reference writes it (simulator.jl:124-133)
    if is_1d(sim)
        VolumeFluxes.current_interior_state(sim)[:] = new_state
    else
        VolumeFluxes.current_interior_state(sim)[1:end, 1:end] = new_state
    end
    # the library applies the boundary condition (update_bc!, simulator.jl:134)
    upload_interior!(sim, box)
    return nothing
end

# set_current_state!(sim, ::Volume) (simulator.jl:137-139): without this method the
# two set_current_state! signatures above and the reference's would be ambiguous.
VolumeFluxes.set_current_state!(sim::VolumeFluxes.Simulator{<:HIPBackend}, new_state::VolumeFluxes.Volume) =
    VolumeFluxes.set_current_state!(sim, VolumeFluxes.InteriorVolume(new_state))

function upload_interior!(sim, box)
    g = sim.grid.ghostcells[1]
    if is_1d(sim)
        # the 1D interior on every interior row of the embedding
        inner = Array(device_full(sim)[g+1:end-g, g+1:end-g, :])
    else
        data = sim.substep_outputs[1]._data
        inner = Array(data[g+1:end-g, g+1:end-g, :])
    end
    check(box.h, ccall((:sfv_set_state, libsinfvm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int32),
                       box.h, inner, SFV_LAYOUT_INTERIOR))
    box.host_valid = false          # ghosts now differ from the mirror until the next read
end

# ---- ensembles (sfv_batch_*) ---------------------------------------------------------
function check_batch(b, rc)
    rc == SFV_OK && return nothing
    msg = unsafe_string(ccall((:sfv_batch_last_error, libsinfvm), Cstring, (Ptr{Cvoid},), b))
    rc == SFV_ERR_ARGUMENT && throw(ArgumentError(msg))
    rc == SFV_ERR_NONFINITE_DT && (@warn msg; return nothing)
    error("libsinfvm_hip: $msg (status $rc)")
end

# simulate_to_time (src/simulator.jl:179-212, without callbacks) for several HIP
# simulators at once: one persistent kernel launch per RK stage covers all of
# them (a replayed HIP graph), each with its own dt until its own endtime.  Each
# ends bitwise where it would alone.  `endtimes` is one value or one per
# simulator.  Members use RungeKutta2 with WallBC or PeriodicBC.  Returns the
# steps of each member.
function simulate_ensemble!(sims::AbstractVector, endtimes; match_endtime=true, maximum_timestep=nothing)
    n = length(sims)
    n > 0 || throw(ArgumentError("an ensemble needs at least one simulator"))
    te = endtimes isa Number ? fill(Float64(endtimes), n) : collect(Float64, endtimes)
    length(te) == n || throw(DimensionMismatch("give one endtime per simulator"))
    boxes = [handle!(s) for s in sims]
    hs = Ptr{Cvoid}[b.h for b in boxes]
    bref = Ref{Ptr{Cvoid}}(C_NULL)
    check_batch(C_NULL, ccall((:sfv_batch_create, libsinfvm), Cint, (Ptr{Ptr{Cvoid}}, Int32, Ref{Ptr{Cvoid}}),
                              hs, n, bref))
    try
        t = Float64[s.t[1] for s in sims]
        steps = zeros(Int64, n)
        mts = maximum_timestep === nothing ? 0.0 : Float64(maximum_timestep)
        rc = ccall((:sfv_batch_run, libsinfvm), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Int32, Float64, Int64, Ptr{Float64}, Ptr{Int64}),
                   bref[], te, match_endtime ? 1 : 0, mts, 0, t, steps)
        for (i, (s, box)) in enumerate(zip(sims, boxes))
            s.t[1] = t[i]
            if steps[i] > 0
                dt = Ref{Float64}(0.0)
                check(box.h, ccall((:sfv_current_timestep, libsinfvm), Cint, (Ptr{Cvoid}, Ref{Float64}), box.h, dt))
                s.current_timestep[1] = dt[]
            end
            box.host_valid = false
        end
        check_batch(bref[], rc)
        return steps
    finally
        ccall((:sfv_batch_destroy, libsinfvm), Cvoid, (Ptr{Cvoid},), bref[])
    end
end

# ---- segment height of the march kernels (autotuned on the first step) -----------
set_march_rows!(sim::VolumeFluxes.Simulator{<:HIPBackend}, rows::Integer) =
    check(handle!(sim).h, ccall((:sfv_set_march_rows, libsinfvm), Cint, (Ptr{Cvoid}, Int32), handle!(sim).h, rows))
march_rows(sim::VolumeFluxes.Simulator{<:HIPBackend}) =
    Int(ccall((:sfv_get_march_rows, libsinfvm), Int32, (Ptr{Cvoid},), handle!(sim).h))

# ---- reverse-mode gradient (sfv_adjoint) ------------------------------------------
# J = sum((h_T - h_target).^2) over the interior after nsteps perform_step!(sim, Inf)
# from the simulator's current state; returns (J, dJ/dB over the corners, dJ/dCz,
# dJ/dU0 over the interior).  The simulator ends at the final state, as after
# nsteps steps of its own.
function adjoint_gradient(sim::VolumeFluxes.Simulator{<:HIPBackend}, nsteps::Integer, h_target::AbstractMatrix;
                          checkpoint_every::Integer=20)
    box = handle!(sim)
    nx, ny = sim.grid.totalcells[1] - 2sim.grid.ghostcells[1], sim.grid.totalcells[2] - 2sim.grid.ghostcells[2]
    tgt = collect(Float64, h_target)
    size(tgt) == (nx, ny) || throw(DimensionMismatch("h_target must be ($nx, $ny)"))
    dB = zeros(Float64, sim.grid.totalcells[1] + 1, sim.grid.totalcells[2] + 1)
    dU0 = zeros(Float64, nx, ny, 3)
    J, dCz, t = Ref{Float64}(0.0), Ref{Float64}(0.0), Ref{Float64}(sim.t[1])
    check(box.h, ccall((:sfv_adjoint, libsinfvm), Cint,
                       (Ptr{Cvoid}, Int64, Int32, Ref{Float64}, Ptr{Float64}, Ref{Float64}, Ptr{Float64},
                        Ref{Float64}, Ptr{Float64}),
                       box.h, nsteps, checkpoint_every, t, tgt, J, dB, dCz, dU0))
    sim.t[1] = t[]
    box.host_valid = false
    return J[], dB, dCz[], dU0
end

# ChainRules rule for the loss of a run, so Zygote / ChainRules users get the
# device adjoint instead of differentiating the stepping loop.  Parameters are
# the corner topography B and Manning's n (Cz = n^2, ImplicitFriction's
# default pattern, src/friction.jl:26); `make_sim(B, n)` builds a fresh
# HIP simulator with its initial state set.  The pullback returns the
# cotangents of B and n (no cotangent for make_sim, nsteps or h_target).
function run_loss(make_sim, B, n, nsteps, h_target)
    J, _, _, _ = adjoint_gradient(make_sim(B, n), nsteps, h_target)
    return J
end

if isdefined(@__MODULE__, :ChainRulesCore) || Base.find_package("ChainRulesCore") !== nothing
    @eval import ChainRulesCore
    @eval function ChainRulesCore.rrule(::typeof(run_loss), make_sim, B, n, nsteps, h_target)
        J, dB, dCz, _ = adjoint_gradient(make_sim(B, n), nsteps, h_target)
        dn = 2n * dCz
        pullback(ȳ) = (ChainRulesCore.NoTangent(), ChainRulesCore.NoTangent(), ȳ .* dB, ȳ * dn,
                       ChainRulesCore.NoTangent(), ChainRulesCore.NoTangent())
        return J, pullback
    end
end

end # module
