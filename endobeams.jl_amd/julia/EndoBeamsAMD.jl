# EndoBeamsAMD.jl — Julia binding of libendobeams_hip.so (include/endobeams_hip.h).
#
# Drop-in for the per-Newton-iteration path of EndoBeams.jl: `run_simulation_gpu!` has the
# signature and control flow of run_simulation! (src/run_simulation.jl:1-119) but hands
# solve_step_dynamics!, update_converged! and update_not_converged! to the device. The node and
# beam structures are read from the caller's BeamsConfiguration exactly as EndoBeams stores them
# (StructArray fields, SVector/SMatrix layouts, 1-based indices), so no reshaping is needed.
module EndoBeamsAMD

using LinearAlgebra

const LIB = joinpath(@__DIR__, "..", "lib", "libendobeams_hip.so")

const EB_ABI_VERSION = Int32(1)
const EB_SURFACE_NONE, EB_SURFACE_SPHERE, EB_SURFACE_SDF = Int32(0), Int32(1), Int32(2)

# field ids (enum eb_field)
const EB_F_U, EB_F_UDOT, EB_F_UDDOT, EB_F_W, EB_F_WDOT, EB_F_WDDOT, EB_F_R, EB_F_DR = 0:7
const EB_F_U_N, EB_F_UDOT_N, EB_F_UDDOT_N, EB_F_W_N, EB_F_WDOT_N, EB_F_WDDOT_N, EB_F_R_N, EB_F_DR_N = 8:15
const EB_F_FEXT, EB_F_FEXT_N = 16, 20

# mirrors eb_model_desc field by field (isbits, C layout)
struct ModelDesc
    abi_version::Int32
    precision::Int32
    n_nodes::Int64
    n_beams::Int64
    X0::Ptr{Float64}
    conn::Ptr{Int64}
    l0::Ptr{Float64}
    Re0::Ptr{Float64}
    props::Ptr{Float64}
    n_fixed::Int64
    fixed_dofs::Ptr{Int64}
    n_gauss::Int32
    pad0::Int32
    gauss_w::Ptr{Float64}
    gauss_z::Ptr{Float64}
    surface::Int32
    sdf_prefiltered::Int32
    kn::Float64
    mu::Float64
    eps_t::Float64
    eta_n::Float64
    n_contact_beams::Int64
    contact_beams::Ptr{Int64}
    sphere_center::NTuple{3,Float64}
    sphere_radius::Float64
    sdf_n::NTuple{3,Int64}
    sdf_d::NTuple{3,Float64}
    sdf_x0::NTuple{3,Float64}
    sdf_samples::Ptr{Float64}
end

struct NewtonParams
    alpha::Float64
    beta::Float64
    gamma::Float64
    tol_res::Float64
    tol_disp::Float64
    max_iterations::Int32
    krylov_maxit::Int32
    krylov_tol::Float64
end

mutable struct Device
    ctx::Ptr{Cvoid}
    keep::Vector{Any}      # arrays the descriptor points at, kept alive for eb_create
end

check(st, ctx, what) = st == 0 || error("$what failed ($st): " * unsafe_string(ccall((:eb_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx)))

flatten(v::AbstractVector) = collect(reinterpret(Float64, collect(v)))   # SVector/SMatrix → Float64

"""props record of the ABI (EB_NPROPS = 14) from a BeamMaterialProperties."""
function beam_record(p)
    K̄ū, K̄Θ̅, K̄Θ̅Θ̅ = p.K̄ⁱⁿᵗ
    J = diag(p.Jᵨ)
    return [p.radius, p.E, K̄ū, diag(K̄Θ̅)..., diag(K̄Θ̅Θ̅)..., J..., p.Aᵨ, p.damping]
end

"""Descriptor of one structure (an EndoBeams BeamsConfiguration + optional RigidInteraction) and
the arrays it points at, which must stay alive until eb_create/eb_create_batch has returned."""
function model_desc(conf, params, inter=nothing; precision::Integer=0)
    nodes, beams = conf.nodes, conf.beams
    X0 = flatten(nodes.X₀)
    conn = Int64[]; for b in beams; push!(conn, b.node1, b.node2); end
    l0 = collect(Float64, beams.l₀)
    Re0 = flatten(beams.Rₑ⁰)
    props = reduce(vcat, [beam_record(b.properties) for b in beams])
    fixed = collect(Int64, conf.bcs.fixed_dofs)
    gw, gz = collect(params.ωᴳ_beams), collect(params.zᴳ_beams)
    keep = Any[X0, conn, l0, Re0, props, fixed, gw, gz]
    surface, kn, mu, epst, etan = EB_SURFACE_NONE, 0.0, 0.0, 0.0, 0.0
    cb = Int64[]; c = (0.0, 0.0, 0.0); R = 0.0
    sdfn, sdfd, sdfx0, samples = (0, 0, 0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), Float64[]
    if inter !== nothing
        pr = inter.properties
        kn, mu, epst, etan = pr.kₙ, pr.μ, pr.εᵗ, pr.ηₙ
        cb = collect(Int64, inter.slave.contact_beams)
        m = inter.master
        if hasfield(typeof(m), :center)
            surface, c, R = EB_SURFACE_SPHERE, Tuple(m.center), m.radius
        else
            # DiscreteSignedDistanceField: hand over the interpolant's prefiltered, padded
            # coefficients (def_interactions.jl:77) with the grid of read_vtk_sdf
            surface = EB_SURFACE_SDF
            samples = collect(Float64, vec(parent(m.sitp.itp.coefs)))
            sdfn = size(m.sitp.itp) .|> Int64
            sdfd = (m.dx, m.dy, m.dz)
            sdfx0 = (m.dom[1], m.dom[3], m.dom[5])
        end
    end
    push!(keep, cb, samples)
    d = ModelDesc(EB_ABI_VERSION, Int32(precision), length(nodes), length(beams), pointer(X0), pointer(conn),
                  pointer(l0), pointer(Re0), pointer(props), length(fixed), pointer(fixed), Int32(length(gw)),
                  Int32(0), pointer(gw), pointer(gz), surface, Int32(surface == EB_SURFACE_SDF), kn, mu, epst, etan, length(cb),
                  pointer(cb), c, R, sdfn, sdfd, sdfx0, pointer(samples))
    return d, keep
end

node_state(nodes) = ((EB_F_U, nodes.u), (EB_F_UDOT, nodes.u̇), (EB_F_UDDOT, nodes.ü), (EB_F_W, nodes.w),
                     (EB_F_WDOT, nodes.ẇ), (EB_F_WDDOT, nodes.ẅ), (EB_F_R, nodes.R), (EB_F_DR, nodes.ΔR))

"""Creates the device context from an EndoBeams BeamsConfiguration (+ optional RigidInteraction)."""
function Device(conf, params, inter=nothing; precision::Integer=0)
    d, keep = model_desc(conf, params, inter; precision=precision)
    ref = Ref{Ptr{Cvoid}}(C_NULL)
    st = GC.@preserve keep ccall((:eb_create, LIB), Cint, (Ref{ModelDesc}, Ref{Ptr{Cvoid}}), d, ref)
    st == 0 || error("eb_create failed ($st)")
    dev = Device(ref[], keep)
    finalizer(x -> ccall((:eb_destroy, LIB), Cint, (Ptr{Cvoid},), x.ctx), dev)
    # NodesBeams state: current and step-n copies, ΔR = I
    for (f, v) in node_state(conf.nodes)
        set_field!(dev, f, flatten(v)); set_field!(dev, f + 8, flatten(v))
    end
    return dev
end

"""One context for many structures (eb_create_batch): `instances` is a vector of
(conf, inter) pairs sharing `params`. Fields address the instances one after the other."""
function DeviceBatch(instances, params; precision::Integer=0)
    built = [model_desc(conf, params, inter; precision=precision) for (conf, inter) in instances]
    descs = [b[1] for b in built]
    keep = Any[descs, [b[2] for b in built]]
    ref = Ref{Ptr{Cvoid}}(C_NULL)
    st = GC.@preserve keep ccall((:eb_create_batch, LIB), Cint, (Ptr{ModelDesc}, Int64, Ref{Ptr{Cvoid}}),
                                 descs, length(descs), ref)
    st == 0 || error("eb_create_batch failed ($st)")
    dev = Device(ref[], keep)
    finalizer(x -> ccall((:eb_destroy, LIB), Cint, (Ptr{Cvoid},), x.ctx), dev)
    for (f, _) in node_state(instances[1][1].nodes)
        v = reduce(vcat, [flatten(Dict(node_state(conf.nodes))[f]) for (conf, _) in instances])
        set_field!(dev, f, v); set_field!(dev, f + 8, v)
    end
    return dev
end

function set_field!(dev::Device, f::Integer, v::Vector{Float64})
    st = GC.@preserve v ccall((:eb_set_field, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64), dev.ctx, f, v, length(v))
    check(st, dev.ctx, "eb_set_field")
end

function get_field(dev::Device, f::Integer, n::Integer)
    v = Vector{Float64}(undef, n)
    st = GC.@preserve v ccall((:eb_get_field, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64), dev.ctx, f, v, n)
    check(st, dev.ctx, "eb_get_field")
    return v
end

newton(p) = NewtonParams(p.α, p.β, p.γ, p.tolerance_residual, p.tolerance_displacement, Int32(p.max_iterations), Int32(200), 1e-11)

"""solve_step_dynamics!(conf, state, params, inter, tⁿ⁺¹, Δt, solver) (src/solve_step_dynamics.jl:1-14)."""
"""update_loads! and update_boundary_conditions! (src/utils_bcs_and_loads.jl:42-64) at time tⁿ⁺¹."""
function update_loads_and_bcs!(dev::Device, conf, tⁿ⁺¹)
    if conf.loads !== nothing && conf.loads.concentrated_force !== nothing
        dofs = collect(Int64, conf.loads.concentrated_force.loaded_dofs)
        vals = [conf.loads.concentrated_force.force_function(tⁿ⁺¹, i) for i in dofs]
        st = GC.@preserve dofs vals ccall((:eb_update_loads, LIB), Cint, (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Float64}),
                                          dev.ctx, length(dofs), dofs, vals)
        check(st, dev.ctx, "eb_update_loads")
    end
    bcs = conf.bcs
    if !isempty(bcs.displaced_indices)
        for i in bcs.displaced_indices
            bcs.imposed_displacements[i] = bcs.imposed_displacements_function(tⁿ⁺¹)
        end
        v = bcs.imposed_displacements
        st = GC.@preserve v ccall((:eb_update_boundary_conditions, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64),
                                  dev.ctx, v, length(v))
        check(st, dev.ctx, "eb_update_boundary_conditions")
    end
    return nothing
end

function solve_step_dynamics!(dev::Device, conf, params, tⁿ⁺¹, Δt)
    update_loads_and_bcs!(dev, conf, tⁿ⁺¹)
    k = Ref{Cint}(0)
    st = ccall((:eb_solve_step_dynamics, LIB), Cint, (Ptr{Cvoid}, Ref{NewtonParams}, Float64, Ref{Cint}),
               dev.ctx, newton(params), Δt, k)
    check(st, dev.ctx, "eb_solve_step_dynamics")
    return Int(k[])
end

update_converged!(dev::Device) = check(ccall((:eb_update_converged, LIB), Cint, (Ptr{Cvoid},), dev.ctx), dev.ctx, "eb_update_converged")
update_not_converged!(dev::Device) = check(ccall((:eb_update_not_converged, LIB), Cint, (Ptr{Cvoid},), dev.ctx), dev.ctx, "eb_update_not_converged")

# execution mode and launch shapes (no counterpart in the reference)
const EB_EXEC_EAGER = 0
const EB_EXEC_GRAPH = 1
const EB_EXEC_FUSED = 2
"""Newton iterations as replays of a captured HIP graph (`graph = true`) or launched one by one."""
set_exec_mode!(dev, graph::Bool) = set_exec_mode!(dev, graph ? EB_EXEC_GRAPH : EB_EXEC_EAGER)
"""`EB_EXEC_EAGER`, `EB_EXEC_GRAPH` or `EB_EXEC_FUSED` (one small structure: each Newton iteration
as the element kernel and one single-workgroup launch; refused for contexts it does not cover)."""
set_exec_mode!(dev, mode::Integer) =
    check(ccall((:eb_set_exec_mode, LIB), Cint, (Ptr{Cvoid}, Cint), dev.ctx, Cint(mode)), dev.ctx, "eb_set_exec_mode")
"""Times the candidate launch shapes on the live context (after a time step) and keeps the
fastest; returns (chain LDS budget in KB, chain-solve LDS mode, gather workgroup size)."""
function autotune!(dev)
    choice = zeros(Int32, 3)
    check(ccall((:eb_autotune, LIB), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Float64}), dev.ctx, choice, C_NULL),
          dev.ctx, "eb_autotune")
    return Tuple(choice)
end
"""Longest run of predictions that skip a learned start that keeps losing (0: evaluate always)."""
set_learned_backoff!(dev, max_skip::Integer) =
    check(ccall((:eb_set_learned_backoff, LIB), Cint, (Ptr{Cvoid}, Int32), dev.ctx, max_skip), dev.ctx,
          "eb_set_learned_backoff")

# the time loop inside the library (eb_run_steps)
struct TimeParams
    initial_timestep::Float64
    min_timestep::Float64
    max_timestep::Float64
    simulation_end_time::Float64
    accelerate_after_success_it::Int32
    stop_on_energy_threshold::Int32
    energy_threshold::Float64
end
mutable struct RunState
    t::Float64
    dt::Float64
    successive::Int32
    stopped::Int32
    steps::Int64
    failed_steps::Int64
    newton_iterations::Int64
    last_k::Int32
    pad::Int32
end
RunState(params) = RunState(0.0, params.initial_timestep, 0, 0, 0, 0, 0, 0, 0)

# loads and imposed displacements of time t, called by the library before every step attempt
function step_callback(ctx::Ptr{Cvoid}, t::Float64, user::Ptr{Cvoid})::Cint
    conf = unsafe_pointer_to_objref(user)
    dev = Device(ctx, Any[])
    try
        update_loads_and_bcs!(dev, conf, t)
        return Cint(0)
    catch
        return Cint(1)
    end
end

"""run_simulation!'s time loop (src/run_simulation.jl:38-112) inside the library: up to
`max_steps` more converged steps from `state`."""
function run_steps!(dev, conf, params, state::RunState, max_steps::Integer)
    tp = TimeParams(params.initial_timestep, params.min_timestep, params.max_timestep, params.simulation_end_time,
                    params.accelerate_after_success_it, params.stop_on_energy_threshold, params.energy_threshold)
    np = newton(params)
    cb = @cfunction(step_callback, Cint, (Ptr{Cvoid}, Float64, Ptr{Cvoid}))
    st = GC.@preserve conf ccall((:eb_run_steps, LIB), Cint,
                                 (Ptr{Cvoid}, Ref{NewtonParams}, Ref{TimeParams}, Int64, Ptr{Cvoid}, Any, Ref{RunState}),
                                 dev.ctx, np, tp, max_steps, cb, conf, state)
    check(st, dev.ctx, "eb_run_steps")
    return state
end

# eb_run_steps_batch: the same loop per batch instance, each with its own time and Δt; states[s]
# is instance s's record (RunState's fields in an isbits struct, so a Vector is contiguous)
struct RunStateRec
    t::Float64
    dt::Float64
    successive::Int32
    stopped::Int32
    steps::Int64
    failed_steps::Int64
    newton_iterations::Int64
    last_k::Int32
    pad::Int32
end
RunStateRec(params) = RunStateRec(0.0, params.initial_timestep, 0, 0, 0, 0, 0, 0, 0)

"""run_simulation!'s time loop (src/run_simulation.jl:38-112) for every instance of a batch with
its own time and Δt, until no running instance has fewer than `max_steps` new converged steps.
Constant loads by default; for time-dependent ones pass
`cb = @cfunction(f, Cint, (Ptr{Cvoid}, Int64, Float64, Ptr{Cvoid}))`, called with the instance
and its time before each of its attempts."""
function run_steps_batch!(dev, params, states::Vector{RunStateRec}, max_steps::Integer;
                          cb::Ptr{Cvoid}=C_NULL, user::Ptr{Cvoid}=C_NULL)
    tp = TimeParams(params.initial_timestep, params.min_timestep, params.max_timestep, params.simulation_end_time,
                    params.accelerate_after_success_it, 0, params.energy_threshold)
    np = newton(params)
    st = ccall((:eb_run_steps_batch, LIB), Cint,
               (Ptr{Cvoid}, Ref{NewtonParams}, Ref{TimeParams}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{RunStateRec}),
               dev.ctx, np, tp, max_steps, cb, user, states)
    check(st, dev.ctx, "eb_run_steps_batch")
    return states
end

"""recompute_at_gausspts_beams! (src/IO/visualization_beams.jl:132-181) on the device: positions,
displacements, velocities (3 × n per beam), strain and stress at n points along every beam."""
function output_points(dev, nbeams::Integer, n::Integer=10)
    m = nbeams * n
    pos, disp, vel = zeros(3m), zeros(3m), zeros(3m)
    strain, stress = zeros(m), zeros(m)
    check(ccall((:eb_output_points, LIB), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                Ptr{Float64}, Ptr{Float64}), dev.ctx, n, pos, disp, vel, strain, stress), dev.ctx, "eb_output_points")
    return pos, disp, vel, strain, stress
end

"""run_simulation!(conf, params, inter) with the hot path on the GPU (src/run_simulation.jl:1-119)."""
function run_simulation_gpu!(conf, params, inter=nothing)
    dev = Device(conf, params, inter)
    check(ccall((:eb_initialize_state_simulation, LIB), Cint, (Ptr{Cvoid},), dev.ctx), dev.ctx, "initialize")
    tⁿ, Δt, successive = 0.0, params.initial_timestep, 0
    while params.simulation_end_time - tⁿ > 1.0e-10
        tⁿ⁺¹ = tⁿ + Δt
        n_it = solve_step_dynamics!(dev, conf, params, tⁿ⁺¹, Δt)
        if n_it > params.max_iterations
            Δt /= 2
            Δt < params.min_timestep && error("Simulation aborted.")
            update_not_converged!(dev)
            successive = 0
            continue
        end
        successive += 1
        if successive ≥ params.accelerate_after_success_it && n_it <= 8
            Δt = min(Δt * 1.5, params.max_timestep)
            successive = 0
        end
        update_converged!(dev)
        tⁿ = tⁿ⁺¹
    end
    # copy the converged node state back into the caller's StructArray
    N = length(conf.nodes)
    u = reinterpret(eltype(conf.nodes.uⁿ), get_field(dev, EB_F_U_N, 3N))
    conf.nodes.uⁿ .= u
    return dev
end

end # module
