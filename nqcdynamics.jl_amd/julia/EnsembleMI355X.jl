# EnsembleMI355X.jl -- the Julia side of the MI355X ensemble backend.
#
# An ensemble algorithm for SciMLBase.solve(::EnsembleProblem, alg, ::EnsembleMI355X; ...)
# (the call at src/Ensembles/run_dynamics.jl:91-97) that runs every trajectory of a supported
# simulation on one MI355X through libnqcd_mi355x.so (include/nqcd_mi355x.h):
#
#   Simulation{FSSH} / Simulation{Ehrenfest}                      BABwithTsit5
#   Simulation{AdiabaticIESH} / Simulation{EhrenfestNA}           VerletwithElectronics
#   RingPolymerSimulation{Classical}                              BCB
#   RingPolymerSimulation{FSSH} / RingPolymerSimulation{Ehrenfest} BCBwithTsit5
#   RingPolymerSimulation{AdiabaticIESH} / {EhrenfestNA}          BCBWavefunction
#   Simulation{CME} / Simulation{BCME}                            Tsit5 at fixed dt
#   Simulation{Classical}                                         VelocityVerlet
#   Simulation{SpinMappingW} / Simulation{eCMM}                   MInt
#   RingPolymerSimulation{NRPMD} / RingPolymerSimulation{eCMM}    RingPolymerMInt
#
# Not executed in this repository (no Julia toolchain in the build image); the same C ABI is
# exercised by the Python mirror (nqcd_mi355x/) and its tests.  Load it next to NQCDynamics:
#
#     include("EnsembleMI355X.jl"); using .EnsembleMI355XBackend
#     run_dynamics(sim, tspan, distribution; output, reduction=MeanReduction(),
#                  ensemble_algorithm=EnsembleMI355X(), trajectories=10^6, dt=1.0)
#
# The Tully models map to device models automatically.  For the other device models (spin-boson,
# double well with Gaussian coupling, harmonic, Newns-Anderson, diatom-surface, MLP surrogate)
# pass the parameter vector of include/nqcd_mi355x.h explicitly:
#     EnsembleMI355X(device_model=(MODEL_NEWNS_ANDERSON, [m, omega, g, dG, Gamma, a, b, M]))
module EnsembleMI355XBackend

using SciMLBase
using NQCDynamics
using NQCDynamics: DynamicsMethods, DynamicsUtils, Ensembles, RingPolymers
using NQCDynamics.DynamicsMethods: SurfaceHoppingMethods, EhrenfestMethods, ClassicalMethods, IntegrationAlgorithms,
      MappingVariableMethods
using NQCModels
using NQCBase: NQCBase
using NQCDistributions: NQCDistributions
using Distributions: Distributions

export EnsembleMI355X, DeviceTermination

const libnqcd = get(ENV, "NQCD_MI355X_LIB", joinpath(@__DIR__, "..", "libnqcd_mi355x.so"))

# --- codes of include/nqcd_mi355x.h ------------------------------------------------------------
const METHOD_FSSH, METHOD_EHRENFEST, METHOD_IESH, METHOD_EHRENFEST_NA, METHOD_RPMD, METHOD_RPSH, METHOD_RP_EHRENFEST,
      METHOD_CME, METHOD_BCME, METHOD_CLASSICAL, METHOD_SPIN_MAPPING, METHOD_ECMM, METHOD_NRPMD = Int32.(0:12)
const ALG_BAB_TSIT5, ALG_VERLET_ELECTRONICS, ALG_BCB, ALG_BCB_TSIT5, ALG_BCB_WAVEFUNCTION, ALG_TSIT5,
      ALG_VELOCITY_VERLET, ALG_MINT, ALG_RP_MINT = Int32.(0:8)
const MODEL_TULLY1, MODEL_TULLY2, MODEL_TULLY3, MODEL_SPIN_BOSON, MODEL_DW_GAUSSIAN, MODEL_HARMONIC,
      MODEL_NEWNS_ANDERSON, MODEL_DIATOM_SURFACE, MODEL_MLP_SURROGATE, MODEL_MORSE_N, MODEL_ANDERSON_HOLSTEIN_1 = Int32.(0:10)
const FIELD_R, FIELD_V, FIELD_SIGMA_RE, FIELD_SIGMA_IM, FIELD_STATE, FIELD_STEPS, FIELD_NHOPS,
      FIELD_SAVE_POP_AD, FIELD_SAVE_POP_DI, FIELD_SAVE_STATE, FIELD_SAVE_CENTROID = Int32.(0:10)
const FIELD_SAVE_R, FIELD_SAVE_EKIN, FIELD_SAVE_EPOT, FIELD_FINAL_POP_DI, FIELD_SAVE_V, FIELD_MAP_X, FIELD_MAP_P,
      FIELD_INIT_POP_DI = Int32(12), Int32(13), Int32(14), Int32(15), Int32(16), Int32(17), Int32(18), Int32(19)
const OUT_ADIABATIC_POPULATION, OUT_DIABATIC_POPULATION, OUT_SCATTERING_1D, OUT_POPULATION_CORRELATION,
      OUT_FINAL_ADIABATIC_POPULATION, OUT_CENTROID_POSITION, OUT_KINETIC_ENERGY, OUT_TOTAL_ENERGY,
      OUT_SCATTERING_1D_DIABATIC, OUT_POSITION, OUT_VELOCITY = Int32.(0:10)
const SAMPLER_GAUSSIAN, SAMPLER_HARMONIC_WIGNER = Int32(0), Int32(2)
const RED_MEAN, RED_SUM = Int32(0), Int32(1)
const CELL_NONE, CELL_BOUNDARY, CELL_TERMINATE = Int32(0), Int32(1), Int32(2)

struct EnsembleMI355X <: SciMLBase.EnsembleAlgorithm
    device::Int32
    persistent::Bool        # persistent work-queue kernel (false: one launch per step)
    fp32_nuclei::Bool       # FP32 nuclei / FP64 electronics (false: FP64 throughout)
    seed::UInt64            # Philox key of the hop and sampling streams
    device_model::Any       # nothing, or (kind::Integer, params::Vector{Float64})
    autotune::Int           # > 0: time the persistent kernel's launch shapes on this many steps first
end
EnsembleMI355X(; device=0, persistent=true, fp32_nuclei=false, seed=0x4E5143444D493335, device_model=nothing,
               autotune=0) =
    EnsembleMI355X(device, persistent, fp32_nuclei, seed, device_model, autotune)

"""
    DeviceTermination(predicate, dof, value)

A `TerminatingCallback` whose predicate the device evaluates after every step
(DynamicsUtils/callbacks.jl:29): `predicate` is `:abs_gt` (|x_dof| > value), `:gt` (x_dof > value) or
`:lt` (x_dof < value), `dof` 1-based in (ndofs, natoms) order.  Pass it as `callback=` to
`run_dynamics` in place of `TerminatingCallback(func)`, whose Julia closure the device cannot run.
"""
struct DeviceTermination
    predicate::Symbol
    dof::Int
    value::Float64
end

# --- C structs (layout of include/nqcd_mi355x.h) -------------------------------------------
struct ProblemDesc
    method::Int32; algorithm::Int32; n_states::Int32; ndofs::Int32; natoms::Int32; nbeads::Int32
    n_electrons::Int32; rescaling::Int32; precision::Int32; schedule::Int32; electronic_substeps::Int32
    decoherence::Int32; masses::Ptr{Float64}; dt::Float64; temperature::Float64; edc_C::Float64
    disable_estimate::Int32; reserved0::Int32; bandwidth::Float64; fermi_level::Float64; mapping_gamma::Float64
end
struct TermDesc
    kind::Int32; dof::Int32; value::Float64
end

struct SamplerDesc
    kind::Int32; initial_state::Int32; state_basis::Int32; n_models::Int32; ntraj::Int64
    params::Ptr{Float64}; nparams::Int64
end

struct TuneResult
    block::Int32; workgroups_per_cu::Int32; n_candidates::Int32; family::Int32; best_ms::Float64; default_ms::Float64
end

check(h, rc) = rc == 0 || error("nqcd_mi355x: ", unsafe_string(ccall((:nqcd_mi355x_last_error, libnqcd), Cstring, (Ptr{Cvoid},), h)))

# --- simulation -> problem ----------------------------------------------------------------
unsupported(x) = throw(ArgumentError("EnsembleMI355X does not support $(typeof(x))"))

method_code(sim::NQCDynamics.Simulation) = method_code(sim.method)
method_code(::SurfaceHoppingMethods.FSSH) = METHOD_FSSH
method_code(::EhrenfestMethods.Ehrenfest) = METHOD_EHRENFEST
method_code(::SurfaceHoppingMethods.AdiabaticIESH) = METHOD_IESH
method_code(::EhrenfestMethods.EhrenfestNA) = METHOD_EHRENFEST_NA
method_code(::SurfaceHoppingMethods.CME) = METHOD_CME
method_code(::SurfaceHoppingMethods.BCME) = METHOD_BCME
method_code(::ClassicalMethods.Classical) = METHOD_CLASSICAL      # Simulation{Classical}: VelocityVerlet
method_code(::MappingVariableMethods.SpinMappingW) = METHOD_SPIN_MAPPING   # Simulation{SpinMappingW}: MInt
method_code(::MappingVariableMethods.eCMM) = METHOD_ECMM                  # Simulation{eCMM}: MInt
method_code(m) = unsupported(m)
function method_code(sim::NQCDynamics.RingPolymerSimulation)
    m = sim.method
    m isa ClassicalMethods.Classical && return METHOD_RPMD
    m isa SurfaceHoppingMethods.FSSH && return METHOD_RPSH
    m isa EhrenfestMethods.Ehrenfest && return METHOD_RP_EHRENFEST
    m isa SurfaceHoppingMethods.AdiabaticIESH && return METHOD_IESH          # with BCBWavefunction
    m isa EhrenfestMethods.EhrenfestNA && return METHOD_EHRENFEST_NA
    m isa MappingVariableMethods.NRPMD && return METHOD_NRPMD                 # with RingPolymerMInt
    m isa MappingVariableMethods.eCMM && return METHOD_ECMM                   # with RingPolymerMInt
    unsupported(m)
end

algorithm_code(::IntegrationAlgorithms.BABwithTsit5) = ALG_BAB_TSIT5
algorithm_code(::IntegrationAlgorithms.VerletwithElectronics) = ALG_VERLET_ELECTRONICS
algorithm_code(::IntegrationAlgorithms.BCB) = ALG_BCB
algorithm_code(::IntegrationAlgorithms.BCBwithTsit5) = ALG_BCB_TSIT5
algorithm_code(::IntegrationAlgorithms.BCBWavefunction) = ALG_BCB_WAVEFUNCTION
algorithm_code(a) = unsupported(a)

# surface_hopping.jl:65,80-91
rescaling_code(m) = hasproperty(m, :rescaling) ? Int32(findfirst(==(m.rescaling), (:standard, :vinversion, :off)) - 1) : Int32(0)

model_code(m::NQCModels.TullyModelOne) = (MODEL_TULLY1, Float64[m.A, m.B, m.C, m.D])
model_code(m::NQCModels.TullyModelTwo) = (MODEL_TULLY2, Float64[m.A, m.B, m.C, m.D, m.E0])
model_code(m::NQCModels.TullyModelThree) = (MODEL_TULLY3, Float64[m.A, m.B, m.C])
model_code(m) = throw(ArgumentError("no automatic device model for $(typeof(m)): pass " *
                                    "EnsembleMI355X(device_model=(kind, params)) (include/nqcd_mi355x.h)"))

function problem_desc(sim, alg, dt, ensemblealg, masses)
    method = method_code(sim)
    ie = method in (METHOD_IESH, METHOD_EHRENFEST_NA)
    rp = sim isa NQCDynamics.RingPolymerSimulation
    dec, C = Int32(0), 0.1
    if method == METHOD_IESH && sim.method.decoherence isa SurfaceHoppingMethods.DecoherenceCorrectionEDC
        dec, C = Int32(1), Float64(sim.method.decoherence.C)   # decoherence_corrections.jl:10-19
    end
    est = method == METHOD_IESH && !sim.method.estimate_probability ? Int32(1) : Int32(0)
    # CME / BCME (cme.jl:63-182): the device integrates with Tsit5 at the fixed step dt whatever
    # solver the call names; the Fermi level is the model's (cme.jl:31, 144).
    cme = method in (METHOD_CME, METHOD_BCME)
    W = method == METHOD_BCME ? Float64(sim.method.bandwidth) : 0.0
    mu = cme ? Float64(NQCModels.fermilevel(sim)) : 0.0
    # Simulation{Classical}: VelocityVerlet (classical.jl:86)
    # Simulation{SpinMappingW | eCMM}: MInt (IntegrationAlgorithms.jl:78-79); RingPolymerSimulation{NRPMD | eCMM}:
    # RingPolymerMInt (:102-103)
    mapping = method in (METHOD_SPIN_MAPPING, METHOD_ECMM, METHOD_NRPMD)
    algc = method == METHOD_CLASSICAL ? ALG_VELOCITY_VERLET : mapping ? (rp ? ALG_RP_MINT : ALG_MINT) :
           cme ? ALG_TSIT5 : algorithm_code(alg)
    γmap = method in (METHOD_ECMM, METHOD_NRPMD) ? Float64(sim.method.γ) : 0.0
    ProblemDesc(method, algc, NQCModels.nstates(sim), size(sim)[1], size(sim)[2],
        rp ? RingPolymers.nbeads(sim) : 1, ie ? NQCModels.nelectrons(sim) : 0, cme ? Int32(2) : rescaling_code(sim.method),
        ensemblealg.fp32_nuclei ? 1 : 0, ensemblealg.persistent ? 0 : 1, 5, dec, pointer(masses), dt,
        Float64(NQCDynamics.get_temperature(sim)), C, est, Int32(0), W, mu, γmap)
end

# --- initial conditions from the Selection closure (selections.jl:38-42,70-73) --------------
# A ProductDistribution whose electronic part is a FermiDiracState (Adiabatic, or Diabatic for
# IESH): the nuclear part is sampled as the Selection samples it, and the occupations (and, for the
# Diabatic form, the orbitals) are drawn on the device (nqcd_mi355x_sample_fermi_dirac) instead of
# one sample_fermi_dirac_distribution per trajectory on the host (iesh.jl:99-183,
# ehrenfest_na.jl:43-70).
function fermi_dirac_selection(prob, sim)
    method_code(sim) in (METHOD_IESH, METHOD_EHRENFEST_NA) || return nothing
    sel = prob.prob_func
    (sel isa Ensembles.RandomSelection || sel isa Ensembles.OrderedSelection) || return nothing
    d = sel.distribution
    d isa NQCDistributions.ProductDistribution || return nothing
    if d.electronic isa NQCDistributions.FermiDiracState{NQCDistributions.Diabatic}
        method_code(sim) == METHOD_IESH || return nothing
    else
        d.electronic isa NQCDistributions.FermiDiracState{NQCDistributions.Adiabatic} || return nothing
    end
    NQCModels.fermilevel(sim) ≈ d.electronic.fermi_level ||
        error("Fermi level of model and distribution do not match")      # iesh.jl:103-112
    return sel
end

function load_fermi_dirac!(h, sel, sim, trajectories, seed)
    d = sel.distribution
    nuc = sel isa Ensembles.OrderedSelection ? [d.nuclear[sel.indices[i]] for i in 1:trajectories] :
          [rand(d.nuclear) for _ in 1:trajectories]
    r = reduce(hcat, [vec(u.r) for u in nuc]); v = reduce(hcat, [vec(u.v) for u in nuc])
    ne = NQCModels.nelectrons(sim)
    st = repeat(Int32.(0:ne - 1), 1, trajectories)
    check(h, ccall((:nqcd_mi355x_load_initial, libnqcd), Int32,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}),
        h, trajectories, r, v, C_NULL, C_NULL, st, C_NULL))
    av = d.electronic.available_states
    avail = av isa Colon ? Int32[] : Int32.(collect(DynamicsUtils.get_available_states(av, NQCModels.nstates(sim)))) .- Int32(1)
    basis = d.electronic isa NQCDistributions.FermiDiracState{NQCDistributions.Diabatic} ? Int32(1) : Int32(0)
    check(h, ccall((:nqcd_mi355x_sample_fermi_dirac, libnqcd), Int32,
        (Ptr{Cvoid}, Float64, Ptr{Int32}, Int32, Int32, UInt64),
        h, Float64(d.electronic.β), isempty(avail) ? C_NULL : avail, length(avail), basis, seed))
end

# --- device sampling of recognised distributions (RandomSelection, selections.jl:70-99) ---------
# A RandomSelection of DynamicalDistribution(velocity, position, size) * PureState(state, basis) whose
# velocities and positions are fixed numbers or Distributions.Normal (what PositionHarmonicWigner /
# VelocityHarmonicWigner give), per degree of freedom or one for all, is sampled on the device
# (nqcd_mi355x_sample_initial, Philox stream 0): the same independent Gaussian law per degree of
# freedom as rand(distribution.nuclear) (selections.jl:95-99), drawn from a different generator,
# instead of one host prob_func call per trajectory.  Lists of configurations and other
# distributions keep the host path.
gaussian_moments(x::Real, n) = (fill(Float64(x), n), zeros(n))
gaussian_moments(x::Distributions.Normal, n) = (fill(Float64(Distributions.mean(x)), n), fill(Float64(Distributions.std(x)), n))
function gaussian_moments(x::AbstractArray, n)
    length(x) == n || return nothing
    m = map(vec(x)) do e
        e isa Real ? (Float64(e), 0.0) :
        e isa Distributions.Normal ? (Float64(Distributions.mean(e)), Float64(Distributions.std(e))) : nothing
    end
    any(isnothing, m) && return nothing
    return (first.(m), last.(m))
end
gaussian_moments(x, n) = nothing

function device_sampler(prob, sim, model_kind)
    sel = prob.prob_func
    sel isa Ensembles.RandomSelection || return nothing
    d = sel.distribution
    d isa NQCDistributions.ProductDistribution || return nothing
    el = d.electronic
    el isa NQCDistributions.PureState || return nothing
    mc = method_code(sim)
    mc in (METHOD_FSSH, METHOD_EHRENFEST, METHOD_CME, METHOD_BCME, METHOD_SPIN_MAPPING, METHOD_ECMM) || return nothing
    sim isa NQCDynamics.RingPolymerSimulation && return nothing   # ring polymers upload their bead configurations
    nuc = d.nuclear
    n = prod(size(sim))
    mv = gaussian_moments(nuc.velocity, n); mr = gaussian_moments(nuc.position, n)
    (mv === nothing || mr === nothing) && return nothing
    basis = el isa NQCDistributions.PureState{NQCDistributions.Diabatic} ? Int32(1) : Int32(0)
    if mc in (METHOD_CME, METHOD_BCME)     # the diabatic state is the discrete state (cme.jl:20-24)
        basis == 1 || throw(ArgumentError("CME / BCME start from a PureState in the Diabatic basis"))
        return (SAMPLER_GAUSSIAN, Float64[mr[1][1], mr[2][1], mv[1][1], mv[2][1]], Int32(el.state - 1), Int32(0))
    end
    if mc == METHOD_SPIN_MAPPING           # mapping variables drawn on the device in the run (spin_mapping.jl:15-33)
        basis == 1 || throw(ArgumentError("spin mapping starts from a PureState in the Diabatic basis"))
        return (SAMPLER_GAUSSIAN, Float64[mr[1][1], mr[2][1], mv[1][1], mv[2][1]], Int32(el.state - 1), Int32(1))
    end
    if mc == METHOD_ECMM                   # any electronic distribution: the sphere draw ignores it (cmm.jl:26)
        return (SAMPLER_GAUSSIAN, Float64[mr[1][1], mr[2][1], mv[1][1], mv[2][1]], Int32(0), Int32(1))
    end
    if model_kind == MODEL_SPIN_BOSON      # params: mu_r(nb), sd_r(nb), mu_v(nb), sd_v(nb)
        return (SAMPLER_HARMONIC_WIGNER, vcat(mr[1], mr[2], mv[1], mv[2]), Int32(el.state - 1), basis)
    end
    model_kind == MODEL_DIATOM_SURFACE && basis == 1 && return nothing   # adiabatic start only
    prm = vec(permutedims(hcat(mr[1], mr[2], mv[1], mv[2])))             # [mu_r, sd_r, mu_v, sd_v] per dof
    return (SAMPLER_GAUSSIAN, prm, Int32(el.state - 1), basis)
end

function load_initial!(h, prob, sim, trajectories, seed, model_kind)
    sel = fermi_dirac_selection(prob, sim)
    sel === nothing || return load_fermi_dirac!(h, sel, sim, trajectories, seed)
    ds = device_sampler(prob, sim, model_kind)
    if ds !== nothing
        kind, prm, state, basis = ds
        sd = Ref(SamplerDesc(kind, state, basis, Int32(1), Int64(trajectories), pointer(prm), length(prm)))
        GC.@preserve prm check(h, ccall((:nqcd_mi355x_sample_initial, libnqcd), Int32,
            (Ptr{Cvoid}, Ref{SamplerDesc}, UInt64), h, sd, seed))
        return
    end
    us = [prob.prob_func(prob.prob, i, false).u0 for i in 1:trajectories]
    r = reduce(hcat, [vec(DynamicsUtils.get_positions(u)) for u in us])
    v = reduce(hcat, [vec(DynamicsUtils.get_velocities(u)) for u in us])
    if method_code(sim) in (METHOD_RPMD, METHOD_CLASSICAL)   # positions and velocities only
        check(h, ccall((:nqcd_mi355x_load_initial, libnqcd), Int32,
            (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}),
            h, trajectories, r, v, C_NULL, C_NULL, C_NULL, C_NULL))
        return
    end
    # (v, r, qmap, pmap): spin_mapping.jl:15-33, cmm.jl:26-34; ring polymers (F, nbeads) qmap / pmap, nrpmd.jl:49-65,
    # rpecmm.jl:11-25
    if method_code(sim) in (METHOD_SPIN_MAPPING, METHOD_ECMM, METHOD_NRPMD)
        n = NQCModels.nstates(sim)
        X = reduce(hcat, [vec(MappingVariableMethods.get_mapping_positions(u)) for u in us])
        P = reduce(hcat, [vec(MappingVariableMethods.get_mapping_momenta(u)) for u in us])
        st = zeros(Int32, trajectories)   # the occupied state only seeds device draws; the variables are given
        check(h, ccall((:nqcd_mi355x_load_initial, libnqcd), Int32,
            (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}),
            h, trajectories, r, v, C_NULL, C_NULL, st, C_NULL))
        check(h, ccall((:nqcd_mi355x_load_mapping, libnqcd), Int32, (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}),
            h, trajectories, X, P))
        return
    end
    if method_code(sim) in (METHOD_CME, METHOD_BCME)   # SurfaceHoppingVariables(r, v, state) only (cme.jl:20-24)
        st = Int32[round(Int32, first(u.state)) - Int32(1) for u in us]
        check(h, ccall((:nqcd_mi355x_load_initial, libnqcd), Int32,
            (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}),
            h, trajectories, r, v, C_NULL, C_NULL, st, C_NULL))
        return
    end
    # density matrix (n, n) for FSSH / Ehrenfest, wavefunction psi (n, n_electrons) for IESH /
    # EhrenfestNA; the discrete state is the active state or the occupied orbitals (1-based)
    q = [DynamicsUtils.get_quantum_subsystem(u) for u in us]
    σr = cat([real.(Matrix(x)) for x in q]...; dims=3); σi = cat([imag.(Matrix(x)) for x in q]...; dims=3)
    st = reduce(hcat, [Int32.(round.(Int, vec(u.state))) .- Int32(1) for u in us])
    check(h, ccall((:nqcd_mi355x_load_initial, libnqcd), Int32,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}),
        h, trajectories, r, v, σr, σi, st, C_NULL))
end

download(h, field, T, dims...) = (a = Array{T}(undef, dims...);
    check(h, ccall((:nqcd_mi355x_download, libnqcd), Int32, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Int64), h, field, a, sizeof(a))); a)
device_reduce(h, kind, red, len) = (a = Vector{Float64}(undef, len);
    check(h, ccall((:nqcd_mi355x_reduce, libnqcd), Int32, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Int64), h, kind, red, a, len)); a)

# --- outputs (DynamicsOutputs.jl, TimeCorrelationFunctions.jl) from the device fields ---------
output_name(f) = f isa Function ? nameof(f) : nameof(typeof(f))
device_output_kind(f) = Dict(:OutputAdiabaticPopulation => OUT_ADIABATIC_POPULATION,
    :OutputDiabaticPopulation => OUT_DIABATIC_POPULATION, :OutputCentroidPosition => OUT_CENTROID_POSITION,
    :OutputStateResolvedScattering1D => OUT_SCATTERING_1D, :PopulationCorrelationFunction => OUT_POPULATION_CORRELATION,
    :OutputKineticEnergy => OUT_KINETIC_ENERGY, :OutputTotalEnergy => OUT_TOTAL_ENERGY,
    :OutputPosition => OUT_POSITION, :OutputVelocity => OUT_VELOCITY)[output_name(f)]

struct DeviceResults
    n::Int; nsave::Int; sz::Tuple
    r::Array{Float64}; v::Array{Float64}; state::Array{Int32}; steps::Vector{Int32}; nhops::Vector{Int32}
    pop_ad::Array{Float64,3}; pop_di::Array{Float64,3}; save_state::Matrix{Int32}
    ekin::Matrix{Float64}; epot::Matrix{Float64}   # empty unless the outputs ask for energies
    final_ad::Matrix{Float64}                       # (n, ntraj) adiabatic populations of the final state
    final_di::Matrix{Float64}                       # (n, ntraj) diabatic populations of the final state, or empty
    save_r::Array{Float64,3}                        # (ndofs * natoms, nsave, ntraj) for CME / BCME / classical, or empty
    save_v::Array{Float64,3}                        # (ndofs * natoms, nsave, ntraj) for classical dynamics, or empty
    init_pop::Matrix{Float64}                       # (n, ntraj) mapping methods' initial estimator, or empty
    corr_norm::Float64                              # normalisation of the population correlation (n for eCMM)
end

statetype_is_adiabatic(f) = hasproperty(f, :statetype) && f.statetype isa NQCDistributions.Adiabatic

function per_trajectory(f, d::DeviceResults, i, sim, t0, dt)
    name = output_name(f)
    name === :OutputAdiabaticPopulation && return [d.pop_ad[:, s, i] for s in 1:d.nsave]
    name === :OutputDiabaticPopulation && return [d.pop_di[:, s, i] for s in 1:d.nsave]
    name === :OutputDiscreteState && return [Int(d.save_state[s, i]) + 1 for s in 1:d.nsave]
    name === :OutputSurfaceHops && return Int(d.nhops[i])
    name === :OutputKineticEnergy && return d.ekin[:, i]                     # DynamicsOutputs.jl:97
    name === :OutputTotalEnergy && return d.ekin[:, i] .+ d.epot[:, i]       # DynamicsOutputs.jl:90
    name === :OutputFinalTime && return t0 + d.steps[i] * dt
    if name === :OutputPosition                               # DynamicsOutputs.jl:62-66
        isempty(d.save_r) && throw(ArgumentError("EnsembleMI355X: positions per save exist for CME / BCME / classical"))
        return [reshape(d.save_r[:, s, i], d.sz) for s in 1:d.nsave]
    end
    if name === :OutputVelocity                               # DynamicsOutputs.jl:35-39
        isempty(d.save_v) && throw(ArgumentError("EnsembleMI355X: velocities per save exist for classical dynamics"))
        return [reshape(d.save_v[:, s, i], d.sz) for s in 1:d.nsave]
    end
    if name === :PopulationCorrelationFunction            # TimeCorrelationFunctions.jl:26-40,61-88
        P = statetype_is_adiabatic(f) ? d.pop_ad : d.pop_di
        K0 = isempty(d.init_pop) || statetype_is_adiabatic(f) ? P[:, 1, i] : d.init_pop[:, i]
        return [d.corr_norm .* (K0 * P[:, s, i]') for s in 1:d.nsave]
    end
    if name === :OutputStateResolvedScattering1D            # DynamicsOutputs.jl:313-338, at the final state
        if f.type === :adiabatic
            p = d.final_ad[:, i]
        elseif f.type === :diabatic
            isempty(d.final_di) && throw(ArgumentError("EnsembleMI355X: diabatic final populations exist for " *
                                                       "two-state one-dof FSSH / Ehrenfest"))
            p = d.final_di[:, i]
        else
            throw(ArgumentError("$(f.type) not recognised. Only `:diabatic` or `:adiabatic` accepted."))
        end
        z = zero(p)
        return first(reshape(d.r[:, i], d.sz)) > 0 ? (reflection=z, transmission=p) : (reflection=p, transmission=z)
    end
    throw(ArgumentError("EnsembleMI355X has no device value for output $name"))
end

# Save schedule of SciML's saveat (forwarded by run_dynamics.jl:91-97): nothing saves every step; a
# number Δ saves t0:Δ:t1; a range or vector must start at t0 and be evenly spaced by a multiple of dt.
# Returns (save_every, nsave, times); the device saves slot k at step k * save_every.
function save_schedule(saveat, tspan, dt)
    nsteps = round(Int, (tspan[2] - tspan[1]) / dt)
    tol = 1e-9 * max(1.0, abs(dt))
    multiple(Δ) = (k = round(Int, Δ / dt); (k >= 1 && abs(k * dt - Δ) <= tol * k) ? k :
        throw(ArgumentError("EnsembleMI355X: saveat spacing $Δ is not a positive multiple of dt = $dt")))
    if saveat === nothing
        k, nsave = 1, nsteps + 1
    elseif saveat isa Number
        k = multiple(Float64(saveat))
        nsteps % k == 0 || throw(ArgumentError("EnsembleMI355X: t1 - t0 must be a multiple of saveat"))
        nsave = nsteps ÷ k + 1
    else
        ts = collect(Float64, saveat)
        isempty(ts) && throw(ArgumentError("EnsembleMI355X: empty saveat"))
        abs(ts[1] - tspan[1]) <= tol || throw(ArgumentError("EnsembleMI355X: saveat must start at tspan[1]"))
        if length(ts) == 1
            return 0, 1, ts
        end
        Δ = diff(ts)
        all(x -> abs(x - Δ[1]) <= tol * length(ts), Δ) ||
            throw(ArgumentError("EnsembleMI355X: saveat times must be evenly spaced (non-uniform saveat is not supported)"))
        k = multiple(Δ[1])
        (length(ts) - 1) * k <= nsteps || throw(ArgumentError("EnsembleMI355X: saveat extends past tspan[2]"))
        return k, length(ts), ts
    end
    return k, nsave, collect(tspan[1] .+ dt * k .* (0:nsave - 1))
end

# --- the ensemble algorithm ----------------------------------------------------------------
function SciMLBase.__solve(prob::SciMLBase.EnsembleProblem, alg, ensemblealg::EnsembleMI355X;
        trajectories, dt=1.0, saveat=nothing, callback=nothing, kwargs...)
    sim = prob.prob.p
    sim isa NQCDynamics.AbstractSimulation || throw(ArgumentError("EnsembleMI355X needs a Simulation"))
    isempty(kwargs) || throw(ArgumentError("EnsembleMI355X does not take $(keys(kwargs))"))
    outputs = prob.output_func.functions       # EnsembleSaver (Ensembles.jl:39-63)
    savetime = prob.output_func.savetime
    tspan = prob.prob.tspan
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(C_NULL, ccall((:nqcd_mi355x_open, libnqcd), Int32, (Int32, Ref{Ptr{Cvoid}}), ensemblealg.device, h))
    try
        masses = collect(Float64, sim.atoms.masses)
        d = problem_desc(sim, alg, dt, ensemblealg, masses)
        GC.@preserve masses check(h[], ccall((:nqcd_mi355x_set_problem, libnqcd), Int32, (Ptr{Cvoid}, Ref{ProblemDesc}), h[], d))
        kind, params = ensemblealg.device_model === nothing ? model_code(sim.cache.model) : ensemblealg.device_model
        check(h[], ccall((:nqcd_mi355x_set_model, libnqcd), Int32, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Int64),
            h[], 0, kind, collect(Float64, params), length(params)))
        load_initial!(h[], prob, sim, trajectories, ensemblealg.seed, kind)
        save_every, nsave, times = save_schedule(saveat, tspan, dt)
        # TerminatingCallback with a predicate the device evaluates (DeviceTermination) and the reference's
        # CellBoundaryCallback / TerminateCellCallback (callbacks.jl:4-22), alone or in a CallbackSet
        cbs = callback === nothing ? () :
              callback isa SciMLBase.CallbackSet ? (callback.discrete_callbacks..., callback.continuous_callbacks...) :
              (callback,)
        dterm = nothing
        cellmode = CELL_NONE
        for cb in cbs
            if cb isa DeviceTermination
                dterm === nothing || throw(ArgumentError("EnsembleMI355X evaluates one DeviceTermination"))
                dterm = cb
            elseif cb isa SciMLBase.DiscreteCallback && cb.condition === DynamicsUtils.outside_cell &&
                   cb.affect! === DynamicsUtils.enforce_periodicity!
                cellmode = CELL_BOUNDARY
            elseif cb isa SciMLBase.DiscreteCallback && cb.condition === DynamicsUtils.outside_cell &&
                   cb.affect! === SciMLBase.terminate!
                cellmode = CELL_TERMINATE
            else
                throw(ArgumentError("EnsembleMI355X evaluates DeviceTermination, CellBoundaryCallback and " *
                                    "TerminateCellCallback on the device"))
            end
        end
        if cellmode != CELL_NONE && sim.cell isa NQCBase.PeriodicCell   # InfiniteCell: the callbacks never fire
            vec_ = Matrix{Float64}(sim.cell.vectors)
            per = Int32.(collect(sim.cell.periodicity))
            check(h[], ccall((:nqcd_mi355x_set_cell, libnqcd), Int32, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Ptr{Int32}),
                h[], cellmode, Int32(size(vec_, 1)), vec_, per))
        end
        kinds = Dict(:abs_gt => Int32(1), :gt => Int32(2), :lt => Int32(3))
        term = dterm === nothing ? C_NULL :
            Ref(TermDesc(kinds[dterm.predicate], Int32(dterm.dof - 1), dterm.value))
        # launch shape of the persistent kernel (a no-op for the problems it does not cover; the
        # choice stays in the library's table for later runs of the same state count and precision)
        if ensemblealg.autotune > 0 && ensemblealg.persistent
            tr = Ref{TuneResult}()
            t1 = tspan[1] + dt * min(ensemblealg.autotune, round(Int, (tspan[2] - tspan[1]) / dt))
            rc = @threadcall((:nqcd_mi355x_autotune, libnqcd), Int32,
                (Ptr{Cvoid}, Float64, Float64, Ptr{TermDesc}, UInt64, Ptr{TuneResult}),
                h[], tspan[1], t1, term, ensemblealg.seed, tr)
            rc == -2 || check(h[], rc)   # NQCD_ERR_UNSUPPORTED: keep the default shape
        end
        # long device call: do not hold the Julia GC
        rc = @threadcall((:nqcd_mi355x_run, libnqcd), Int32, (Ptr{Cvoid}, Float64, Float64, Int32, Ptr{TermDesc}, UInt64),
            h[], tspan[1], tspan[2], Int32(save_every), term, ensemblealg.seed)
        check(h[], rc)
        n = NQCModels.nstates(sim)
        # the device's slots run to t1; a saveat vector that stops earlier keeps the first nsave
        ndev = save_every == 0 ? 1 : div(round(Int, (tspan[2] - tspan[1]) / dt), save_every) + 1
        red = prob.reduction
        ie = method_code(sim) in (METHOD_IESH, METHOD_EHRENFEST_NA)
        cme = method_code(sim) in (METHOD_CME, METHOD_BCME)
        cls = method_code(sim) == METHOD_CLASSICAL
        spinmap = method_code(sim) in (METHOD_SPIN_MAPPING, METHOD_ECMM, METHOD_NRPMD)
        two_state = !ie && !cme && !cls && !spinmap && !(sim isa NQCDynamics.RingPolymerSimulation) && n == 2
        device_reducible(f) = begin
            nm = output_name(f)
            cls && return nm in (:OutputKineticEnergy, :OutputTotalEnergy, :OutputPosition, :OutputVelocity)
            spinmap && return nm in (:OutputDiabaticPopulation, :OutputKineticEnergy, :OutputTotalEnergy) ||
                              (nm === :PopulationCorrelationFunction && !statetype_is_adiabatic(f)) ||
                              (nm === :OutputCentroidPosition && sim isa NQCDynamics.RingPolymerSimulation)
            cme && return nm in (:OutputKineticEnergy, :OutputPosition) ||
                          (nm === :OutputTotalEnergy && method_code(sim) == METHOD_CME)
            nm in (:OutputAdiabaticPopulation, :OutputDiabaticPopulation, :OutputCentroidPosition,
                   :OutputKineticEnergy, :OutputTotalEnergy) ||
            (nm === :PopulationCorrelationFunction && !statetype_is_adiabatic(f)) ||
            (nm === :OutputStateResolvedScattering1D && (f.type === :adiabatic || (f.type === :diabatic && two_state)))
        end
        if (red isa Ensembles.MeanReduction || red isa Ensembles.SumReduction) && all(device_reducible, outputs)
            # reduced on the device in a fixed order (reductions.jl:8-52)
            rk = red isa Ensembles.MeanReduction ? RED_MEAN : RED_SUM
            out = Ensembles.Dictionary{Symbol,Any}()
            savetime && insert!(out, :Time, times)
            for f in outputs
                k = device_output_kind(f)
                if k == OUT_SCATTERING_1D
                    k = f.type === :diabatic ? OUT_SCATTERING_1D_DIABATIC : k
                    a = device_reduce(h[], k, rk, 2n); insert!(out, output_name(f), (reflection=a[1:n], transmission=a[n+1:2n]))
                elseif k in (OUT_POSITION, OUT_VELOCITY)
                    D = prod(size(sim))
                    a = reshape(device_reduce(h[], k, rk, D * ndev), D, ndev)
                    insert!(out, output_name(f), [reshape(a[:, s], size(sim)) for s in 1:nsave])
                elseif k in (OUT_CENTROID_POSITION, OUT_KINETIC_ENERGY, OUT_TOTAL_ENERGY)
                    insert!(out, output_name(f), device_reduce(h[], k, rk, ndev)[1:nsave])
                elseif k == OUT_POPULATION_CORRELATION
                    a = reshape(device_reduce(h[], k, rk, n * n * ndev), n, n, ndev)
                    insert!(out, output_name(f), [a[:, :, s] for s in 1:nsave])
                else
                    a = reshape(device_reduce(h[], k, rk, n * ndev), n, ndev)
                    insert!(out, output_name(f), [a[:, s] for s in 1:nsave])
                end
            end
            return SciMLBase.EnsembleSolution([out], 0.0, true)
        end
        # every other reduction: the per-trajectory dictionaries EnsembleSaver would have made,
        # handed to the problem's own reduction functor (Sort*, FileReduction, ...)
        nst = ie ? d.n_electrons : 1
        wants_energy = any(f -> output_name(f) in (:OutputKineticEnergy, :OutputTotalEnergy), outputs)
        state = cls || spinmap ? zeros(Int32, nst, trajectories) : download(h[], FIELD_STATE, Int32, nst, trajectories)
        # adiabatic populations of the final state: one-hot of the active state (FSSH, fssh.jl:144-148),
        # occupied orbitals (IESH, iesh.jl:370-374), diag(sigma) (Ehrenfest, ehrenfest.jl:70-73)
        final_ad = zeros(n, trajectories)
        if d.method in (METHOD_FSSH, METHOD_RPSH, METHOD_IESH, METHOD_CME, METHOD_BCME)
            for i in 1:trajectories, e in 1:nst
                final_ad[state[e, i] + 1, i] = 1.0
            end
        elseif d.method in (METHOD_EHRENFEST, METHOD_RP_EHRENFEST)
            σ = download(h[], FIELD_SIGMA_RE, Float64, n, n, trajectories)
            for i in 1:trajectories, a in 1:n
                final_ad[a, i] = σ[a, a, i]
            end
        end
        final_di = two_state &&
                   any(f -> output_name(f) === :OutputStateResolvedScattering1D && f.type === :diabatic, outputs) ?
                   download(h[], FIELD_FINAL_POP_DI, Float64, n, trajectories) : zeros(0, 0)
        elec = d.method != METHOD_RPMD && !cme && !cls && !spinmap  # RPMD, CME / BCME, classical, spin mapping: no density
        res = DeviceResults(n, nsave, size(sim),
            download(h[], FIELD_R, Float64, prod(size(sim)), trajectories), download(h[], FIELD_V, Float64, prod(size(sim)), trajectories),
            state, download(h[], FIELD_STEPS, Int32, trajectories),
            cls || spinmap ? zeros(Int32, trajectories) : download(h[], FIELD_NHOPS, Int32, trajectories),
            elec ? download(h[], FIELD_SAVE_POP_AD, Float64, n, ndev, trajectories)[:, 1:nsave, :] : zeros(n, nsave, trajectories),
            elec || spinmap ? download(h[], FIELD_SAVE_POP_DI, Float64, n, ndev, trajectories)[:, 1:nsave, :] : zeros(n, nsave, trajectories),
            ie || !(elec || cme) ? zeros(Int32, nsave, trajectories) : download(h[], FIELD_SAVE_STATE, Int32, ndev, trajectories)[1:nsave, :],
            wants_energy ? download(h[], FIELD_SAVE_EKIN, Float64, ndev, trajectories)[1:nsave, :] : zeros(0, 0),
            wants_energy ? download(h[], FIELD_SAVE_EPOT, Float64, ndev, trajectories)[1:nsave, :] : zeros(0, 0),
            final_ad, final_di,
            (cme || cls) && any(f -> output_name(f) === :OutputPosition, outputs) ?
                download(h[], FIELD_SAVE_R, Float64, prod(size(sim)), ndev, trajectories)[:, 1:nsave, :] : zeros(0, 0, 0),
            cls && any(f -> output_name(f) === :OutputVelocity, outputs) ?
                download(h[], FIELD_SAVE_V, Float64, prod(size(sim)), ndev, trajectories)[:, 1:nsave, :] : zeros(0, 0, 0),
            spinmap ? download(h[], FIELD_INIT_POP_DI, Float64, n, trajectories) : zeros(0, 0),
            d.method == METHOD_ECMM ? Float64(n) : 1.0)
        batch = map(1:trajectories) do i
            out = Ensembles.Dictionary{Symbol,Any}()
            savetime && insert!(out, :Time, times)
            for f in outputs
                insert!(out, output_name(f), per_trajectory(f, res, i, sim, tspan[1], dt))
            end
            out
        end
        u, _ = red(prob.u_init === nothing ? [] : prob.u_init, batch, 1:trajectories)
        return SciMLBase.EnsembleSolution(u, 0.0, true)
    finally
        ccall((:nqcd_mi355x_close, libnqcd), Cvoid, (Ptr{Cvoid},), h[])
    end
end

end # module
