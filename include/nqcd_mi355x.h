/*
 * nqcd_mi355x.h -- C ABI of libnqcd_mi355x.so, the MI355X (gfx950) ensemble
 * trajectory integrator that backs the EnsembleMI355X ensemble algorithm.
 *
 * The reference (NQCDynamics.jl v1.0.5) has no C ABI: ensemble integration is
 * reached through run_dynamics -> SciMLBase.solve(ensemble_problem, algorithm,
 * ensemble_algorithm; trajectories, ...) (src/Ensembles/run_dynamics.jl:91-97).
 * Each entry point below replaces one piece of what that call does for the
 * supported (method, algorithm, model) combinations; the citation on each
 * declaration names the reference code it stands in for.
 *
 * Conventions
 *  - Status: int32_t, 0 = OK, negative = error class (NQCD_ERR_*); the message
 *    is available from nqcd_mi355x_last_error(handle) until the next call on
 *    that handle.  No C++ exception crosses this boundary.
 *  - Ownership: the caller owns every host buffer for the duration of a call;
 *    the library copies inputs and never retains host pointers.  Device memory
 *    belongs to the handle and is freed by nqcd_mi355x_close.
 *  - Threading: one handle = one HIP stream.  Calls on one handle must be
 *    serialised by the caller and return after the device work completed.
 *  - Layout: host arrays use the reference's column-major Julia layout with the
 *    trajectory index last: positions/velocities (ndofs, natoms[, nbeads], ntraj),
 *    density matrices (nstates, nstates, ntraj) as separate real/imaginary parts.
 *    States are 0-based (the Julia side subtracts 1).
 *  - Determinism: results are a pure function of the inputs and philox_key; they
 *    do not depend on launch configuration or trajectory scheduling.
 */
#ifndef NQCD_MI355X_H
#define NQCD_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NQCD_MI355X_ABI_VERSION 2

typedef struct nqcd_handle nqcd_handle_t;

enum nqcd_status {
    NQCD_OK = 0,
    NQCD_ERR_ARG = -1,          /* invalid argument (shape, null pointer, range) */
    NQCD_ERR_UNSUPPORTED = -2,  /* method/algorithm/model combination not supported */
    NQCD_ERR_DEVICE = -3,       /* HIP runtime error */
    NQCD_ERR_STATE = -4,        /* call out of order (e.g. run before load) */
    NQCD_ERR_NOMEM = -5
};

/* DynamicsMethods (src/DynamicsMethods/): FSSH fssh.jl:23, Ehrenfest ehrenfest.jl:27,
 * AdiabaticIESH iesh.jl:26, EhrenfestNA ehrenfest_na.jl:6, RingPolymerSimulation{Classical}
 * classical.jl:28, RingPolymerSimulation{FSSH} rpsh.jl:8, RingPolymerSimulation{Ehrenfest}
 * ehrenfest_rpmd.jl:3. */
enum nqcd_method {
    NQCD_METHOD_FSSH = 0,
    NQCD_METHOD_EHRENFEST = 1,
    NQCD_METHOD_IESH = 2,
    NQCD_METHOD_EHRENFEST_NA = 3,
    NQCD_METHOD_RPMD = 4,
    NQCD_METHOD_RPSH = 5,
    NQCD_METHOD_RP_EHRENFEST = 6,
    NQCD_METHOD_CME = 7,          /* classical master equation, cme.jl:63-105 */
    NQCD_METHOD_BCME = 8,         /* broadened CME, cme.jl:115-182 (bandwidth in nqcd_problem_desc) */
    NQCD_METHOD_CLASSICAL = 9,    /* Simulation{Classical}, classical.jl (single-state model, one coordinate) */
    NQCD_METHOD_SPIN_MAPPING = 10,/* Simulation{SpinMappingW}, spin_mapping.jl:1-131 (2 to 4 states, one coordinate) */
    NQCD_METHOD_ECMM = 11,        /* Simulation{eCMM}, cmm.jl:1-120 (2 to 4 states, one coordinate; gamma in
                                     nqcd_problem_desc.mapping_gamma); with RP_MINT RingPolymerSimulation{eCMM},
                                     rpecmm.jl:1-103 */
    NQCD_METHOD_NRPMD = 12        /* RingPolymerSimulation{NRPMD}, nrpmd.jl:1-139 (gamma in mapping_gamma; the
                                     reference's default is 0.5) */
};

/* IntegrationAlgorithms (IntegrationAlgorithms.jl:19-56): BABwithTsit5, VerletwithElectronics,
 * BCB, BCBwithTsit5. */
enum nqcd_algorithm {
    NQCD_ALG_BAB_TSIT5 = 0,
    NQCD_ALG_VERLET_ELECTRONICS = 1,
    NQCD_ALG_BCB = 2,
    NQCD_ALG_BCB_TSIT5 = 3,
    NQCD_ALG_BCB_WAVEFUNCTION = 4,   /* ring-polymer IESH / EhrenfestNA (bcb_wavefunction.jl) */
    NQCD_ALG_TSIT5 = 5,              /* the nuclear ODE by Tsit5 at a fixed step (CME / BCME, IntegrationAlgorithms.jl:68) */
    NQCD_ALG_VELOCITY_VERLET = 6,    /* OrdinaryDiffEq's VelocityVerlet (Simulation{Classical}, IntegrationAlgorithms.jl:61) */
    NQCD_ALG_MINT = 7,               /* MInt, the mapping integrator (mint.jl:42-181; select_algorithm, IntegrationAlgorithms.jl:79) */
    NQCD_ALG_RP_MINT = 8             /* RingPolymerMInt (ringpolymer_mint.jl:28-131; IntegrationAlgorithms.jl:102-103):
                                        NRPMD and ring-polymer eCMM, 1 to 64 beads */
};

/* rescaling keyword of Simulation{FSSH} (fssh.jl:41, surface_hopping.jl:219,234-245). */
enum nqcd_rescaling { NQCD_RESCALE_STANDARD = 0, NQCD_RESCALE_VINVERSION = 1, NQCD_RESCALE_OFF = 2 };

/* Nuclear precision: FP64 everywhere (parity path) or FP32 positions, velocities,
 * forces and model evaluation with FP64 eigen/NAC/electronics. */
enum nqcd_precision { NQCD_PREC_FP64 = 0, NQCD_PREC_FP32_NUCLEI = 1 };

/* Scheduling: persistent work-queue kernel (trajectory state in registers) or one
 * launch per nuclear step with the state in HBM (the plain GPU port). */
enum nqcd_schedule { NQCD_SCHED_PERSISTENT = 0, NQCD_SCHED_STEP_PER_LAUNCH = 1 };

/* Model kinds (NQCModels, external to the reference tree). */
enum nqcd_model_kind {
    NQCD_MODEL_TULLY1 = 0,       /* params: A, B, C, D                         */
    NQCD_MODEL_TULLY2 = 1,       /* params: A, B, C, D, E0                     */
    NQCD_MODEL_TULLY3 = 2,       /* params: A, B, C                            */
    NQCD_MODEL_SPIN_BOSON = 3,   /* params: eps, Delta, omega[nb], c[nb]       */
    NQCD_MODEL_DW_GAUSSIAN = 4,  /* params: a, b, offset, C, D                 */
    NQCD_MODEL_HARMONIC = 5,     /* params: m, omega, r0                       */
    /* Newns-Anderson impurity (MiaoSubotnik diabats) with a trapezoidal bath discretisation:
     * params: m, omega, g, dG, Gamma, a, b, M.  U0 = m w^2 x^2/2, impurity h = m w^2 (x-g)^2/2 + dG - U0,
     * bath eps_k = a + k (b-a)/(M-1), coupling sqrt(Gamma (b-a)/(M-1) / 2pi); n_states = M + 1. */
    NQCD_MODEL_NEWNS_ANDERSON = 6,
    /* Diatomic molecule above a surface, 6 coordinates (ndofs = 3, natoms = 2), two diabats:
     * params: Dn an bn Aw bw Di ai bi Ds as Zs Delta C g Zc mu with bond b = |R1 - R2| and
     * height Z = mu z1 + (1 - mu) z2:  V11 = Dn (1 - e^{-an(b-bn)})^2 + Aw e^{-bw Z},
     * V22 = Di (1 - e^{-ai(b-bi)})^2 + Ds (1 - e^{-as(Z-Zs)})^2 + Delta, V12 = C e^{-g (Z-Zc)^2}. */
    NQCD_MODEL_DIATOM_SURFACE = 7,
    /* The MLP surrogate of the diatom-surface model (trained with nqcd_mi355x_surrogate_train):
     * params: fallback threshold (Hartree), then the 16 diatom-surface parameters of the analytic
     * model used wherever the two ensemble members disagree by more than the threshold. */
    NQCD_MODEL_MLP_SURROGATE = 8,
    /* n-state Morse model of one coordinate (Coker and Xiao, J. Chem. Phys. 102, 496 (1995); NQCModels'
     * ThreeStateMorse for n = 3): params: n (2..8), then [D_i, beta_i, r_i, c_i] per state, then
     * [A_ij, a_ij, r_ij] per pair i < j in row-major order:
     * V_ii = D_i (1 - e^{-beta_i (x - r_i)})^2 + c_i,  V_ij = A_ij e^{-a_ij (x - r_ij)^2}.
     * FSSH / Ehrenfest with n_states >= the slot's n; slots of different n (and two-state Tully / double-well
     * slots) mix in one ensemble. */
    NQCD_MODEL_MORSE_N = 9,
    /* One-mode Anderson-Holstein impurity of the CME / BCME tests (test/Dynamics/cme.jl:27-55; Dou, Nitzan,
     * Subotnik, J. Chem. Phys. 142, 084110 (2015)): params: k, Ed, c, V12 with V11 = k x^2/2,
     * V22 = V11 + Ed + c x, V12 constant (hybridisation Gamma = 2 pi V12^2). */
    NQCD_MODEL_ANDERSON_HOLSTEIN_1 = 10
};

/* IESH decoherence keyword (iesh.jl:72-73): none, or Granucci-Persico EDC
 * (decoherence_corrections.jl:10-39). */
enum nqcd_decoherence { NQCD_DECOHERENCE_NONE = 0, NQCD_DECOHERENCE_EDC = 1 };

typedef struct {
    int32_t method;              /* enum nqcd_method */
    int32_t algorithm;           /* enum nqcd_algorithm */
    int32_t n_states;
    int32_t ndofs;
    int32_t natoms;
    int32_t nbeads;              /* 1 for Simulation */
    int32_t n_electrons;         /* IESH / EhrenfestNA */
    int32_t rescaling;           /* enum nqcd_rescaling */
    int32_t precision;           /* enum nqcd_precision */
    int32_t schedule;            /* enum nqcd_schedule */
    int32_t electronic_substeps; /* Tsit5 sub-steps per nuclear step; 0 -> 5 (bab_electronics.jl:42) */
    int32_t decoherence;         /* IESH: enum nqcd_decoherence */
    const double* masses;        /* natoms (simulations.jl:86-91) */
    double dt;                   /* nuclear time step (run_dynamics.jl:71-73 default 1.0) */
    double temperature;          /* ring polymers: omega_n = nbeads * temperature (ring_polymer.jl:18) */
    double edc_C;                /* EDC constant C (Hartree); 0 -> 0.1 (decoherence_corrections.jl:14) */
    int32_t disable_estimate;    /* IESH: 1 = estimate_probability=false (iesh.jl:73,241-244) */
    int32_t reserved0;
    double bandwidth;            /* BCME: bandwidth of the broadening integrals (cme.jl:130-133) */
    double fermi_level;          /* CME / BCME: Fermi level mu of the metal (NQCModels.fermilevel) */
    double mapping_gamma;        /* eCMM / NRPMD: the gamma keyword of Simulation{eCMM}(atoms, model; gamma)
                                    (cmm.jl:22-24) or RingPolymerSimulation{NRPMD | eCMM}(...; gamma) (nrpmd.jl:42,
                                    rpecmm.jl:7); ignored otherwise (spin mapping's gamma follows from n_states) */
} nqcd_problem_desc;

/* TerminatingCallback (DynamicsUtils/callbacks.jl:15-29) with a predicate the device evaluates after
 * each step and not at t0, as a DiscreteCallback's condition (a trajectory that starts past the boundary
 * takes one step): |x_dof| > value, x_dof > value or x_dof < value.  FSSH / Ehrenfest only. */
enum nqcd_term_kind { NQCD_TERM_NONE = 0, NQCD_TERM_ABS_GT = 1, NQCD_TERM_GT = 2, NQCD_TERM_LT = 3 };

/* Cell callbacks (DynamicsUtils/callbacks.jl:4-22) on the simulation cell of nqcd_mi355x_set_cell:
 * CellBoundaryCallback wraps the atoms into the cell after a step in which one left it
 * (enforce_periodicity!), TerminateCellCallback ends the trajectory after such a step.  Both follow the
 * hop callback, as the problem's callbacks precede the user's; a TerminatingCallback predicate is
 * evaluated after the wrap. */
enum nqcd_cell_mode { NQCD_CELL_NONE = 0, NQCD_CELL_BOUNDARY = 1, NQCD_CELL_TERMINATE = 2 };
typedef struct {
    int32_t kind;
    int32_t dof;                 /* flattened (ndofs, natoms) index */
    double value;
} nqcd_term_desc;

/* Device-side initial-condition samplers (stand in for NQCDistributions' rand at
 * src/Ensembles/selections.jl:70-99 for the recognised distributions). */
enum nqcd_sampler_kind {
    /* params = [mu_r, sd_r, mu_v, sd_v] per flattened dof (4*D values). */
    NQCD_SAMPLER_GAUSSIAN = 0,
    /* Mixed Tully ensemble: model_id = traj % n_models, p0 = p_lo + ((traj / n_models) % n_p),
     * x ~ N(x0, p0/(20 sqrt 2)), v = p0/m.  params = [x0, p_lo, n_p]. */
    NQCD_SAMPLER_TULLY_GRID = 1,
    /* Gaussian bath of a system-bath model (PositionHarmonicWigner/VelocityHarmonicWigner
     * widths): params = [mu_r(nb), sd_r(nb), mu_v(nb), sd_v(nb)]; counters as GAUSSIAN. */
    NQCD_SAMPLER_HARMONIC_WIGNER = 2,
    /* Thermal ring polymers in the harmonic approximation of a well: params = [x0, w_well, beta];
     * normal mode k ~ N(sqrt(P) x0 delta_k0, 1/(beta_P m (w_k^2 + w_well^2))), velocities
     * ~ N(0, 1/(beta_P m)), beta_P = beta/P (stands in for the PIMC sampler of
     * src/InitialConditions/MetropolisHastings.jl). */
    NQCD_SAMPLER_RING_POLYMER_THERMAL = 3,
    /* Path-integral Monte Carlo (run_monte_carlo_sampling for a RingPolymerSimulation,
     * MetropolisHastings.jl:131-148,170-258): one Metropolis-Hastings chain per ring polymer started with
     * every bead at x0, `passes` steps of a centroid move (beads shifted by (u - 1/2) Delta / sqrt(P)),
     * P / L normal-mode moves (L = max(1, round(segment P)) modes in 1..P-1 resampled from N(0, sigma_k)),
     * each tested against min(1, exp(-(E' - E)/(P T))) with E the bead-summed potential (element [1] of the
     * model matrix), then the reference's second test of the last proposal; velocities ~ N(0, sqrt(P T/m)).
     * params = [x0, Delta, passes, segment(, width(, reassess))]: width 0 (default) takes the reference's
     * sigma_k = sqrt(w_n/m)/(m w_k^2/2), 1 the free ring polymer's sqrt(w_n/m)/w_k; reassess 1 (default)
     * keeps the second test, 0 drops it (the chain then samples the path-integral distribution exactly
     * with width 1).  Random numbers: Philox stream 4. */
    NQCD_SAMPLER_RING_POLYMER_PIMC = 4
};
typedef struct {
    int32_t kind;
    int32_t initial_state;       /* pure electronic state, 0-based */
    int32_t state_basis;         /* 0 = adiabatic, 1 = diabatic (NQCDistributions PureState) */
    int32_t n_models;            /* models set with nqcd_mi355x_set_model slots 0..n_models-1 */
    int64_t ntraj;
    const double* params;
    int64_t nparams;
} nqcd_sampler_desc;

/* Outputs reduced on the device (DynamicsOutputs.jl / TimeCorrelationFunctions.jl). */
enum nqcd_output_kind {
    NQCD_OUT_ADIABATIC_POPULATION = 0,   /* per save slot, nstates   (DynamicsOutputs.jl:198-203) */
    NQCD_OUT_DIABATIC_POPULATION = 1,    /* per save slot, nstates   (DynamicsOutputs.jl:186-189) */
    NQCD_OUT_SCATTERING_1D = 2,          /* final, [reflection(n), transmission(n)] (DynamicsOutputs.jl:313-338) */
    NQCD_OUT_POPULATION_CORRELATION = 3, /* per save slot, n x n P(0)P(t)^T diabatic (TimeCorrelationFunctions.jl:61-88) */
    NQCD_OUT_FINAL_ADIABATIC_POPULATION = 4, /* final, nstates */
    NQCD_OUT_CENTROID_POSITION = 5,      /* per save slot, ring-polymer centroid (OutputCentroidPosition, DynamicsOutputs.jl:73-79) */
    NQCD_OUT_KINETIC_ENERGY = 6,         /* per save slot (OutputKineticEnergy, DynamicsOutputs.jl:97) */
    NQCD_OUT_TOTAL_ENERGY = 7,           /* per save slot, kinetic + classical_potential_energy (OutputTotalEnergy,
                                            DynamicsOutputs.jl:90; fssh.jl:150-154, ehrenfest.jl:85-95) */
    NQCD_OUT_SCATTERING_1D_DIABATIC = 8, /* final, [reflection(n), transmission(n)] of the diabatic populations
                                            (OutputStateResolvedScattering1D(sim, :diabatic)); two-state one-dof */
    NQCD_OUT_POSITION = 9,               /* per save slot, the coordinate (OutputPosition, DynamicsOutputs.jl:62-66);
                                            CME / BCME / classical */
    NQCD_OUT_VELOCITY = 10               /* per save slot, the velocity (OutputVelocity, DynamicsOutputs.jl:35-39); classical */
};
/* IESH: DIABATIC_POPULATION is iesh_diabatic_population (iesh.jl:337-369), state 0 = impurity;
 * EhrenfestNA (which the reference gives no diabatic estimator) reports sum_e |(U psi)_{a e}|^2. */
/* Reductions (src/Ensembles/reductions.jl:8-52). */
enum nqcd_reduction_kind { NQCD_RED_MEAN = 0, NQCD_RED_SUM = 1 };

/* Per-trajectory fields for nqcd_mi355x_download (SortByTrajectoryReduction data). */
enum nqcd_field {
    NQCD_FIELD_R = 0,            /* double (ndofs, natoms[, nbeads], ntraj) final positions   */
    NQCD_FIELD_V = 1,            /* double, final velocities                                  */
    NQCD_FIELD_SIGMA_RE = 2,     /* double (n, n, ntraj) final density matrix, real part      */
    NQCD_FIELD_SIGMA_IM = 3,     /* double (n, n, ntraj) imaginary part                       */
    NQCD_FIELD_STATE = 4,        /* int32 (ntraj) final active state                          */
    NQCD_FIELD_STEPS = 5,        /* int32 (ntraj) nuclear steps taken                         */
    NQCD_FIELD_NHOPS = 6,        /* int32 (ntraj) accepted hops (OutputSurfaceHops)            */
    NQCD_FIELD_SAVE_POP_AD = 7,  /* double (n, nsave, ntraj) adiabatic population per save     */
    NQCD_FIELD_SAVE_POP_DI = 8,  /* double (n, nsave, ntraj) diabatic population per save      */
    NQCD_FIELD_SAVE_STATE = 9,   /* int32 (nsave, ntraj) active state per save                 */
    NQCD_FIELD_SAVE_CENTROID = 10, /* double (nsave, ntraj) ring-polymer centroid per save      */
    NQCD_FIELD_FALLBACKS = 11,     /* int32 (ntraj) surrogate steps that fell back to the analytic model */
    NQCD_FIELD_SAVE_R = 12,        /* double (ndofs * natoms, nsave, ntraj) configuration per save (diatom-surface) */
    NQCD_FIELD_SAVE_EKIN = 13,     /* double (nsave, ntraj) kinetic energy per save (two-state one-dof FSSH / Ehrenfest) */
    NQCD_FIELD_SAVE_EPOT = 14,     /* double (nsave, ntraj) classical potential energy per save (same problems) */
    NQCD_FIELD_FINAL_POP_DI = 15,  /* double (n, ntraj) diabatic populations of the final state (fssh.jl:132-142,
                                      ehrenfest.jl:75-83); two-state one-dof FSSH / Ehrenfest */
    NQCD_FIELD_SAVE_V = 16,        /* double (nsave, ntraj) velocity per save (classical dynamics) */
    NQCD_FIELD_MAP_X = 17,         /* double (n[, nbeads], ntraj) final mapping positions (u.qmap) */
    NQCD_FIELD_MAP_P = 18,         /* double (n[, nbeads], ntraj) final mapping momenta (u.pmap) */
    NQCD_FIELD_INIT_POP_DI = 19    /* double (n, ntraj) mapping methods: the initial estimator of the population
                                      correlation at t0 (initial_diabatic_population, Estimators.jl:121, cmm.jl:106-110,
                                      rpecmm.jl:95-103) */
};
/* IESH / EhrenfestNA shapes: SIGMA_RE / SIGMA_IM hold the wavefunction psi (n, n_electrons, ntraj),
 * STATE the occupied adiabatic states (n_electrons, ntraj), SAVE_POP_AD / SAVE_POP_DI (n, nsave, ntraj). */

typedef struct {
    double kernel_ms;            /* device time of the trajectory kernels (HIP events) */
    double wall_ms;              /* host wall time of nqcd_mi355x_run */
    int64_t traj_steps;          /* sum over trajectories of nuclear steps taken */
    int64_t launches;            /* kernel launches (or graph replays) issued */
    int32_t block;               /* workgroup size used */
    int32_t grid;                /* workgroups launched */
} nqcd_run_stats;

/* ABI version of the loaded library (compare with NQCD_MI355X_ABI_VERSION). No reference counterpart. */
int32_t nqcd_mi355x_abi_version(void);

/* Create a handle on HIP device `device` (one stream).  Replaces the ensemble_algorithm
 * object handed to SciMLBase.solve (run_dynamics.jl:49,91-97). */
int32_t nqcd_mi355x_open(int32_t device, nqcd_handle_t** out);

/* Release every device allocation of the handle. */
void nqcd_mi355x_close(nqcd_handle_t* h);

/* Describe the simulation: Simulation/RingPolymerSimulation fields read by the integrator
 * (simulations.jl:12-73,86-94), the method (fssh.jl:23-45 etc.) and the algorithm chosen
 * by select_algorithm (IntegrationAlgorithms.jl:65-103). */
int32_t nqcd_mi355x_set_problem(nqcd_handle_t* h, const nqcd_problem_desc* desc);

/* The simulation cell and the cell callback of the following runs (NQCBase's PeriodicCell, the `cell`
 * keyword of Simulation, simulations.jl:12-44; callbacks.jl:4-22).  `vectors` holds the ndofs x ndofs
 * cell matrix column-major with the lattice vectors as columns (PeriodicCell(vectors)), `periodic`
 * ndofs flags (NULL: all periodic).  An atom is outside when a fractional coordinate f = vectors^-1 x
 * is below 0 or above 1 (check_atoms_in_cell); the wrap maps f_j to mod(f_j, 1) on the periodic
 * dimensions of every atom and sets x = vectors f (apply_cell_boundaries!).  mode NQCD_CELL_NONE
 * clears the cell (InfiniteCell).  Supported by the one-coordinate FSSH / Ehrenfest (two and n
 * states), CME / BCME and classical problems (ndofs = 1) and the diatom-surface problem (ndofs = 3,
 * two atoms); set_problem clears it. */
int32_t nqcd_mi355x_set_cell(nqcd_handle_t* h, int32_t mode, int32_t ndofs, const double* vectors, const int32_t* periodic);

/* Set model slot `slot` (0..3).  Replaces the NQCModels model held in sim.cache
 * (Simulation cache, simulations.jl:28-30); several slots let one ensemble mix models,
 * selected per trajectory by model_id. */
int32_t nqcd_mi355x_set_model(nqcd_handle_t* h, int32_t slot, int32_t model_kind, const double* params, int64_t n);

/* Upload initial conditions for ntraj trajectories (the DynamicsVariables built by
 * sample_distribution/prob_func, selections.jl:38-99, fssh.jl:47-65, ehrenfest.jl:43-48).
 * state may be NULL (sampled from diag(sigma), fssh.jl:54) and model_id may be NULL (slot 0).
 * IESH / EhrenfestNA (iesh.jl:89-184, ehrenfest_na.jl:43-58): state holds the occupied adiabatic
 * states (n_electrons, ntraj), sig_re/sig_im the wavefunction (n, n_electrons, ntraj) or both NULL
 * for psi[state_e, e] = 1. */
int32_t nqcd_mi355x_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v,
                                 const double* sig_re, const double* sig_im, const int32_t* state,
                                 const int32_t* model_id);

/* Spin mapping / eCMM: set the mapping positions X and momenta P (n, ntraj) of the loaded trajectories
 * (the qmap / pmap of DynamicsVariables, spin_mapping.jl:15-33, cmm.jl:26-34) instead of drawing them in
 * the first step of nqcd_mi355x_run; X = P = NULL returns to drawing.  Without this call every run draws
 * them from Philox stream 4 of philox_key: spin mapping theta_i = 2 pi u (counter (trajectory, i)),
 * radius sqrt(2 + gamma) on the trajectory's occupied diabatic state and sqrt(gamma) elsewhere,
 * X_i = radius sin(theta_i), P_i = radius cos(theta_i); eCMM a point uniform on the 2n-sphere of radius
 * sqrt(2 + 2 n gamma) (normals (z_2k, z_2k+1) at counter (trajectory, k), normalised; X the first n
 * coordinates, P the last n).  Call after load_initial or sample_initial, with their ntraj. */
int32_t nqcd_mi355x_load_mapping(nqcd_handle_t* h, int64_t ntraj, const double* X, const double* P);

/* Sample initial conditions on the device from a recognised distribution
 * (RandomSelection, selections.jl:70-99) using Philox stream 0 of `seed`. */
int32_t nqcd_mi355x_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* sampler, uint64_t seed);

/* IESH / EhrenfestNA: replace the occupied states of the loaded trajectories by a Fermi-Dirac sample
 * at each trajectory's initial position (DynamicsVariables(sim, v, r, FermiDiracState), iesh.jl:99-183,
 * ehrenfest_na.jl:43-70): the Boltzmann-factor Metropolis chain of sample_fermi_dirac_distribution
 * (DynamicsUtils.jl:191-205), n_available * n_electrons moves from the lowest available states, draws
 * from Philox stream 3 of `seed`.  basis 0 (Adiabatic): the chain weighs the adiabatic energies and
 * psi[state_e, e] = 1.  basis 1 (Diabatic, IESH only, iesh.jl:138-183): the chain weighs the diagonal
 * of the model matrix, each sampled diabatic state d then draws an adiabatic state with weights
 * U_db^2 until a free one comes up, and the electron's orbital is row d of U (sign of U_d0).
 * beta = 1/kT (INFINITY for zero temperature); available_states = NULL makes every state available
 * (get_available_states(:, n), DynamicsUtils.jl:210-214), otherwise it lists 0-based states, taken as
 * a set.  Call after load_initial or sample_initial. */
int32_t nqcd_mi355x_sample_fermi_dirac(nqcd_handle_t* h, double beta, const int32_t* available_states,
                                       int32_t n_available, int32_t basis, uint64_t seed);

/* Integrate every trajectory over [t0, t1] with the method's callbacks (hopping after each
 * step, surface_hopping.jl:2-16; termination, callbacks.jl:15-29) and record the save
 * slots.  save_every > 0 saves slot k at step k*save_every (slot 0 = t0), mirroring
 * saveat = t0:save_every*dt:t1.  Hop uniforms come from Philox stream 1 of philox_key.
 * Replaces SciMLBase.solve(ensemble_problem, algorithm, ensemble_algorithm; ...). */
int32_t nqcd_mi355x_run(nqcd_handle_t* h, double t0, double t1, int32_t save_every,
                        const nqcd_term_desc* term, uint64_t philox_key);

/* Reduce an output over trajectories (MeanReduction / SumReduction, reductions.jl:8-52):
 * out receives len doubles (n per save slot, 2n for SCATTERING_1D, n*n per save slot for
 * POPULATION_CORRELATION, column-major). */
int32_t nqcd_mi355x_reduce(nqcd_handle_t* h, int32_t output_kind, int32_t reduction_kind, double* out, int64_t len);

/* The same reduction with no device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The partial sums and the result go through one 2 MiB scratch array per
 * device (262144 doubles) that the library holds until the process ends, resolved when the first
 * handle is opened on the device; outputs computed from the final state are evaluated on the fly
 * instead of being materialised.  A reduction needing rows x ceil(ntraj / 4096, at most 1024 chunks)
 * (twice that for TOTAL_ENERGY) + rows doubles beyond that limit fails with NQCD_ERR_ARG.  Results are
 * bit-identical to nqcd_mi355x_reduce.  Every call on a device shares that device's array: calls that
 * overlap on one device (from different handles or threads) race. */
int32_t nqcd_mi355x_reduce_noalloc(nqcd_handle_t* h, int32_t output_kind, int32_t reduction_kind, double* out, int64_t len);

/* Copy a per-trajectory field to host memory (SortByTrajectoryReduction / EnsembleSaver data,
 * Ensembles.jl:39-63, reductions.jl:54-55).  bytes must equal the field size. */
int32_t nqcd_mi355x_download(nqcd_handle_t* h, int32_t field, void* out, int64_t bytes);

/* Statistics of the last nqcd_mi355x_run (device time from HIP events on the handle's stream).
 * Replaces the @timed wall clock around solve (run_dynamics.jl:91). */
int32_t nqcd_mi355x_last_run_stats(nqcd_handle_t* h, nqcd_run_stats* out);

typedef struct {
    int32_t block;               /* chosen threads per workgroup */
    int32_t workgroups_per_cu;   /* chosen resident workgroups per compute unit (0 = occupancy limit) */
    int32_t n_candidates;        /* launch shapes timed */
    int32_t family;              /* 1 two-state one-dof kernel, 2 diatom-surface kernel */
    double best_ms;              /* kernel time of the chosen shape on the sample run */
    double default_ms;           /* kernel time of the default shape (256 threads, occupancy limit) */
} nqcd_tune_result;

/* Autotuner of the persistent work-queue kernels (no counterpart in the reference; the natural
 * caller is the one-step precompile_dynamics pre-run of run_dynamics.jl:100-120).  Runs the
 * loaded problem over [t0, t1] (no saves; with `term` when given, which makes the sample include
 * the tail of uneven trajectory lengths) once per candidate launch shape --
 * 256, 128 and 64 threads per workgroup, each at the occupancy limit and at 3/4, 1/2 and 1/4 of it --
 * keeps the fastest (the default unless a candidate is faster by more than 2%) in a process-wide
 * table keyed by kernel family, state count, precision and compute-unit count, and later runs
 * of that key use it.  With NQCD_MI355X_TUNE_CACHE set the table persists in that file.  Results
 * of a run do not depend on the shape.  The sample runs overwrite the handle's results; call
 * nqcd_mi355x_run afterwards.  Two-state one-dof (persistent schedule) and analytic
 * diatom-surface problems only; NQCD_ERR_UNSUPPORTED otherwise. */
int32_t nqcd_mi355x_autotune(nqcd_handle_t* h, double t0, double t1, const nqcd_term_desc* term, uint64_t philox_key,
                             nqcd_tune_result* out);

/* ---- MLP surrogate of the diabatic Hamiltonian (no counterpart in the reference) -------------
 * A 6 -> 128 -> 128 -> 3 tanh network for (H11, H22, H12) of a six-coordinate two-state model,
 * two ensemble members, trained on the GPU with BF16 MFMA forward / backward kernels and Adam. */
typedef struct {
    int32_t steps;               /* optimiser steps per member */
    int32_t batch;               /* samples per step (>= 32) */
    double lr, lr_final;         /* Adam learning rate, decayed geometrically to lr_final */
    double loss_weights[3];      /* weights of the H11, H22, H12 squared errors */
    uint64_t seed;               /* initial weights and minibatch order */
} nqcd_train_desc;

/* Train both members on X (n, 6) -> Y (n, 3) (row-major, Hartree / bohr).  Inputs and outputs are
 * normalised by the data's mean and standard deviation.  loss_history (2 * steps) may be NULL. */
int32_t nqcd_mi355x_surrogate_train(nqcd_handle_t* h, const double* X, const double* Y, int64_t n,
                                    const nqcd_train_desc* desc, double* loss_history);
/* H (n, 3) and dH/dX (n, 3, 6) of one member at X (n, 6); dH may be NULL. */
int32_t nqcd_mi355x_surrogate_eval(nqcd_handle_t* h, int32_t member, const double* X, int64_t n, double* H, double* dH);
/* Gradient of the training loss with respect to the 17795 FP32 parameters of one member on (X, Y),
 * computed by the training kernels without an update (loss weights w[3]). */
int32_t nqcd_mi355x_surrogate_gradient(nqcd_handle_t* h, int32_t member, const double* X, const double* Y, int64_t n,
                                       const double* w, double* grad, double* loss);
/* Parameters of one member: W1 (128x6), b1, W2 (128x128), b2, W3 (3x128), b3 row-major (17795
 * floats); norm = in_shift[6], in_scale[6], out_shift[3], out_scale[3] (may be NULL on set). */
int32_t nqcd_mi355x_surrogate_set_params(nqcd_handle_t* h, int32_t member, const float* params, const float* norm);
int32_t nqcd_mi355x_surrogate_get_params(nqcd_handle_t* h, int32_t member, float* params, float* norm);

/* Test hook (no counterpart in the reference): determinants of nmat complex n x n matrices (re, im
 * row-major, matrix after matrix, n <= 64) by the device routine behind the IESH hopping estimate
 * (the det S of FastDeterminant.jl:20-22 as used at iesh.jl:242-253); out[3 k .. 3 k + 2] = Re det,
 * Im det, 1 (0 when a pivot column is zero). */
int32_t nqcd_mi355x_test_determinant(nqcd_handle_t* h, const double* re, const double* im, int32_t n, int32_t nmat,
                                     double* out);

/* Test hook (no counterpart in the reference): the BCME broadening integrals n(h) = int A f and
 * n2(h) = int (e - h) A f over [mu - W/2, mu + W/2] (evaluate_broadening / evaluate_force_broadening,
 * cme.jl:167-182) by the device routine behind the BCME force, for count values of h. */
int32_t nqcd_mi355x_test_broadening(nqcd_handle_t* h, const double* hs, int64_t count, double mu, double kT, double Gamma,
                                    double W, double* n, double* n2);

/* Message for the last failed call on h (thread-local when h is NULL). */
const char* nqcd_mi355x_last_error(nqcd_handle_t* h);

#ifdef __cplusplus
}
#endif
#endif /* NQCD_MI355X_H */
