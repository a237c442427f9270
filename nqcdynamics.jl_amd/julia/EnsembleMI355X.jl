# EnsembleMI355X.jl -- the Julia side of the MI355X ensemble backend.
#
# An ensemble algorithm for SciMLBase.solve(::EnsembleProblem, alg, ::EnsembleMI355X; ...)
# (the call at src/Ensembles/run_dynamics.jl:91-97) that runs every trajectory of a
# supported Simulation on one MI355X through libnqcd_mi355x.so (include/nqcd_mi355x.h).
#
# Not executed in this repository (no Julia toolchain in the build image); the same
# C ABI is exercised by the Python mirror and tests.  Load it next to NQCDynamics:
#
#     include("EnsembleMI355X.jl"); using .EnsembleMI355XBackend
#     run_dynamics(sim, tspan, distribution; output, reduction=MeanReduction(),
#                  ensemble_algorithm=EnsembleMI355X(), trajectories=10^6, dt=1.0)
module EnsembleMI355XBackend

using SciMLBase
using NQCDynamics
using NQCDynamics: DynamicsMethods, Ensembles
using NQCModels

export EnsembleMI355X

const libnqcd = get(ENV, "NQCD_MI355X_LIB", joinpath(@__DIR__, "..", "libnqcd_mi355x.so"))

struct EnsembleMI355X <: SciMLBase.EnsembleAlgorithm
    device::Int32
    persistent::Bool
    seed::UInt64
end
EnsembleMI355X(; device=0, persistent=true, seed=0x4E5143444D493335) = EnsembleMI355X(device, persistent, seed)

# --- C structs (layout of include/nqcd_mi355x.h) -------------------------------------------
struct ProblemDesc
    method::Int32; algorithm::Int32; n_states::Int32; ndofs::Int32; natoms::Int32; nbeads::Int32
    n_electrons::Int32; rescaling::Int32; precision::Int32; schedule::Int32; electronic_substeps::Int32
    decoherence::Int32; masses::Ptr{Float64}; dt::Float64; temperature::Float64; edc_C::Float64
    disable_estimate::Int32; reserved::NTuple{7,Int32}
end
struct TermDesc
    kind::Int32; dof::Int32; value::Float64
end
struct RunStats
    kernel_ms::Float64; wall_ms::Float64; traj_steps::Int64; launches::Int64; block::Int32; grid::Int32
end

check(h, rc) = rc == 0 || error("nqcd_mi355x: ", unsafe_string(ccall((:nqcd_mi355x_last_error, libnqcd), Cstring, (Ptr{Cvoid},), h)))

method_code(::NQCDynamics.SurfaceHoppingMethods.FSSH) = Int32(0)
method_code(::NQCDynamics.EhrenfestMethods.Ehrenfest) = Int32(1)
method_code(m) = throw(ArgumentError("EnsembleMI355X does not support $(typeof(m))"))
rescaling_code(m) = hasproperty(m, :rescaling) ? Int32(findfirst(==(m.rescaling), (:standard, :vinversion, :off)) - 1) : Int32(0)

model_code(m::NQCModels.TullyModelOne) = (Int32(0), Float64[m.A, m.B, m.C, m.D])
model_code(m::NQCModels.TullyModelTwo) = (Int32(1), Float64[m.A, m.B, m.C, m.D, m.E0])
model_code(m::NQCModels.TullyModelThree) = (Int32(2), Float64[m.A, m.B, m.C])
model_code(m) = throw(ArgumentError("EnsembleMI355X has no device model for $(typeof(m))"))

function SciMLBase.__solve(prob::SciMLBase.EnsembleProblem, alg, ensemblealg::EnsembleMI355X;
        trajectories, dt=1.0, saveat=nothing, callback=nothing, kwargs...)
    sim = prob.prob.p
    sim isa NQCDynamics.Simulation || throw(ArgumentError("EnsembleMI355X needs a Simulation"))
    tspan = prob.prob.tspan
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(C_NULL, ccall((:nqcd_mi355x_open, libnqcd), Int32, (Int32, Ref{Ptr{Cvoid}}), ensemblealg.device, h))
    try
        masses = collect(Float64, sim.atoms.masses)
        d = ProblemDesc(method_code(sim.method), 0, NQCModels.nstates(sim), size(sim)[1], size(sim)[2], 1, 0,
            rescaling_code(sim.method), 0, ensemblealg.persistent ? 0 : 1, 5, 0, pointer(masses), dt, 0.0, 0.1, 0,
            ntuple(_ -> 0, 7))
        GC.@preserve masses check(h[], ccall((:nqcd_mi355x_set_problem, libnqcd), Int32, (Ptr{Cvoid}, Ref{ProblemDesc}), h[], d))
        kind, params = model_code(sim.cache.model)
        check(h[], ccall((:nqcd_mi355x_set_model, libnqcd), Int32, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Int64),
            h[], 0, kind, params, length(params)))
        # initial conditions from the Selection closure (selections.jl:38-42,70-73)
        us = [prob.prob_func(prob.prob, i, 1).u0 for i in 1:trajectories]
        r = reduce(hcat, [vec(u.r) for u in us]); v = reduce(hcat, [vec(u.v) for u in us])
        σr = cat([Matrix(u.σreal) for u in us]...; dims=3); σi = cat([Matrix(u.σimag) for u in us]...; dims=3)
        st = Int32[round(Int32, first(u.state)) - 1 for u in us]
        check(h[], ccall((:nqcd_mi355x_load_initial, libnqcd), Int32,
            (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}),
            h[], trajectories, r, v, σr, σi, st, C_NULL))
        save_every = saveat === nothing ? 0 : round(Int32, step(saveat) / dt)
        term = callback === nothing ? C_NULL : Ref(TermDesc(1, callback.dof - 1, callback.value))
        # long device call: do not hold the Julia GC
        rc = @threadcall((:nqcd_mi355x_run, libnqcd), Int32, (Ptr{Cvoid}, Float64, Float64, Int32, Ptr{TermDesc}, UInt64),
            h[], tspan[1], tspan[2], save_every, term, ensemblealg.seed)
        check(h[], rc)
        n = NQCModels.nstates(sim); nsave = save_every == 0 ? 1 : div(round(Int, (tspan[2] - tspan[1]) / dt), save_every) + 1
        popad = Array{Float64}(undef, n, nsave, trajectories)
        check(h[], ccall((:nqcd_mi355x_download, libnqcd), Int32, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Int64),
            h[], 7, popad, sizeof(popad)))
        rf = similar(r); check(h[], ccall((:nqcd_mi355x_download, libnqcd), Int32, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Int64), h[], 0, rf, sizeof(rf)))
        stf = Vector{Int32}(undef, trajectories)
        check(h[], ccall((:nqcd_mi355x_download, libnqcd), Int32, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Int64), h[], 4, stf, sizeof(stf)))
        # per-trajectory dictionaries reduced with the problem's reduction (Ensembles.jl:39-63)
        batch = [Ensembles.Dictionary{Symbol,Any}([:Time, :OutputFinalState, :OutputFinalPosition],
                    [collect(tspan), stf[i] + 1, rf[:, i]]) for i in 1:trajectories]
        u, _ = prob.reduction(prob.u_init === nothing ? [] : prob.u_init, batch, 1:trajectories)
        return SciMLBase.EnsembleSolution(u, 0.0, true)
    finally
        ccall((:nqcd_mi355x_close, libnqcd), Cvoid, (Ptr{Cvoid},), h[])
    end
end

end # module
