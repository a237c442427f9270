"""run_dynamics with the EnsembleMI355X ensemble algorithm (src/Ensembles/run_dynamics.jl:42-135).

The reference forwards `ensemble_algorithm` to SciMLBase.solve (run_dynamics.jl:91-97);
here EnsembleMI355X is that algorithm: it hands the whole ensemble to
libnqcd_mi355x.so and returns what the given reduction would produce over the
per-trajectory EnsembleSaver dictionaries (Ensembles.jl:39-63).
"""
import logging

import numpy as np

from . import _lib as L
from .distributions import Adiabatic, DynamicalDistribution, FermiDiracState, ProductDistribution, PureState
from .outputs import (FileReduction, MeanReduction, Output, SortByOutputReduction, SortByTrajectoryReduction,
                      SumReduction, output_name)
from .simulations import (AdiabaticIESH, BCME, CME, Classical, Ehrenfest, EhrenfestNA, FSSH, PeriodicCell,
                          RingPolymerSimulation, Simulation, SpinMappingW, eCMM, NRPMD)

log = logging.getLogger("nqcd_mi355x")


class EnsembleMI355X:
    """Ensemble algorithm running every trajectory on one MI355X (device index `device`).

    schedule: "persistent" (work-queue kernel, default) or "step_per_launch" (plain port).
    autotune: time the persistent kernel's launch shapes on the first `autotune_steps` steps of the
    ensemble before the run (FSSH / Ehrenfest on the two-state one-dof and diatom-surface models; a
    no-op elsewhere); the choice is kept for later runs of the same state count and precision.
    """

    def __init__(self, device=0, schedule="persistent", seed=0x4E5143444D493335, autotune=False, autotune_steps=200):
        if schedule not in ("persistent", "step_per_launch"):
            raise ValueError("schedule must be 'persistent' or 'step_per_launch'")
        self.device, self.schedule, self.seed = device, schedule, seed
        self.autotune, self.autotune_steps = bool(autotune), int(autotune_steps)
        self.tuning = None


class TerminatingCallback:
    """TerminatingCallback (DynamicsUtils/callbacks.jl:15-29) with a predicate on one coordinate that
    the device evaluates after every step: "abs_gt" |x_dof| > value (default), "gt" x_dof > value,
    "lt" x_dof < value.  `dof` is 1-based in (ndofs, natoms) order."""

    PREDICATES = {"abs_gt": L.TERM_ABS_GT, "gt": L.TERM_GT, "lt": L.TERM_LT}

    def __init__(self, dof=1, value=10.0, predicate="abs_gt"):
        if predicate not in self.PREDICATES:
            raise ValueError(f"predicate must be one of {sorted(self.PREDICATES)}")
        self.dof, self.value, self.predicate = int(dof), float(value), predicate


class CellBoundaryCallback:
    """CellBoundaryCallback() (DynamicsUtils/callbacks.jl:9-15): after a step in which an atom left the
    simulation's PeriodicCell, wrap the atoms into the cell (evaluated on the device)."""


class TerminateCellCallback:
    """TerminateCellCallback() (DynamicsUtils/callbacks.jl:17-22): end the trajectory after a step in which
    an atom left the simulation's PeriodicCell (evaluated on the device)."""


def _callbacks(callback):
    """(termination predicate, cell mode) of a callback or a tuple / list of them (a CallbackSet)."""
    cbs = [] if callback is None else list(callback) if isinstance(callback, (tuple, list)) else [callback]
    term, cell = None, L.CELL_NONE
    for cb in cbs:
        if isinstance(cb, TerminatingCallback):
            if term is not None:
                raise ValueError("EnsembleMI355X evaluates one TerminatingCallback predicate")
            term = (cb.dof - 1, cb.value, TerminatingCallback.PREDICATES[cb.predicate])
        elif isinstance(cb, (CellBoundaryCallback, TerminateCellCallback)):
            if cell != L.CELL_NONE:
                raise ValueError("EnsembleMI355X evaluates one cell callback")
            cell = L.CELL_BOUNDARY if isinstance(cb, CellBoundaryCallback) else L.CELL_TERMINATE
        else:
            raise ValueError("EnsembleMI355X evaluates TerminatingCallback(|x_dof| > value), CellBoundaryCallback and "
                             "TerminateCellCallback on the device")
    return term, cell


def _kind(sim):
    """'sh' (FSSH/Ehrenfest), 'ie' (IESH/EhrenfestNA, also as ring polymers with BCBWavefunction), 'rp'
    (ring polymers), 'cme' (CME / BCME), 'cl' (classical dynamics) or 'map' (spin mapping); raises for the
    rest."""
    m = sim.method
    if isinstance(sim, RingPolymerSimulation):
        if isinstance(m, (Classical, FSSH, Ehrenfest)):
            return "rp"
        if isinstance(m, (NRPMD, eCMM)) and sim.ndofs * sim.natoms == 1:
            return "rpmap"
        if isinstance(m, (AdiabaticIESH, EhrenfestNA)):
            return "ie"
    elif isinstance(m, (FSSH, Ehrenfest)):
        return "sh"
    elif isinstance(m, (AdiabaticIESH, EhrenfestNA)):
        return "ie"
    elif isinstance(m, (CME, BCME)):
        return "cme"
    elif isinstance(m, Classical) and sim.model.nstates == 1 and sim.ndofs * sim.natoms == 1:
        return "cl"
    elif isinstance(m, (SpinMappingW, eCMM)) and sim.ndofs * sim.natoms == 1:
        return "map"
    raise ValueError(f"EnsembleMI355X does not support {type(sim).__name__}{{{type(m).__name__}}}")


def _problem(E, sim, dt):
    method = sim.method
    kind = _kind(sim)
    common = dict(precision=L.PREC_FP64 if sim.precision == "fp64" else L.PREC_FP32_NUCLEI,
                  schedule=L.SCHED_PERSISTENT if E._schedule == "persistent" else L.SCHED_STEP_PER_LAUNCH,
                  rescaling=getattr(method, "rescaling_code", 0))
    if kind == "sh":
        E.set_problem(method.method_code, sim.model.nstates, sim.ndofs, sim.natoms, sim.atoms.masses, dt,
                      algorithm=method.algorithm, temperature=sim.temperature, **common)
    elif kind == "ie":
        edc = isinstance(method, AdiabaticIESH) and method.decoherence is not None
        rp = isinstance(sim, RingPolymerSimulation)
        E.set_problem(method.method_code, sim.model.nstates, sim.ndofs, sim.natoms, sim.atoms.masses, dt,
                      algorithm=L.ALG_BCB_WAVEFUNCTION if rp else method.algorithm, n_electrons=sim.model.nelectrons,
                      nbeads=sim.nbeads, temperature=sim.temperature,
                      decoherence=L.DECOHERENCE_EDC if edc else L.DECOHERENCE_NONE,
                      edc_C=method.decoherence.C if edc else 0.1,
                      estimate_probability=getattr(method, "estimate_probability", True), **common)
    elif kind == "cl":
        if sim.precision != "fp64":
            raise ValueError("classical dynamics run in FP64")
        E.set_problem(L.METHOD_CLASSICAL, sim.model.nstates, sim.ndofs, sim.natoms, sim.atoms.masses, dt,
                      algorithm=L.ALG_VELOCITY_VERLET, temperature=sim.temperature, **common)
    elif kind == "rpmap":
        if sim.precision != "fp64":
            raise ValueError("NRPMD / ring-polymer eCMM run in FP64")
        E.set_problem(method.method_code, sim.model.nstates, sim.ndofs, sim.natoms, sim.atoms.masses, dt,
                      algorithm=L.ALG_RP_MINT, nbeads=sim.nbeads, temperature=sim.temperature,
                      mapping_gamma=method.gamma, **common)
    elif kind == "map":
        if sim.precision != "fp64":
            raise ValueError("spin mapping / eCMM run in FP64")
        E.set_problem(method.method_code, sim.model.nstates, sim.ndofs, sim.natoms, sim.atoms.masses, dt,
                      algorithm=L.ALG_MINT, temperature=sim.temperature,
                      mapping_gamma=getattr(method, "gamma", 0.0) if isinstance(method, eCMM) else 0.0, **common)
    elif kind == "cme":
        if sim.precision != "fp64":
            raise ValueError("CME / BCME run in FP64")
        E.set_problem(method.method_code, sim.model.nstates, sim.ndofs, sim.natoms, sim.atoms.masses, dt,
                      algorithm=L.ALG_TSIT5, temperature=sim.temperature, bandwidth=method.bandwidth,
                      fermi_level=getattr(sim.model, "fermi_level", 0.0), **common)
    else:
        code = {Classical: L.METHOD_RPMD, FSSH: L.METHOD_RPSH, Ehrenfest: L.METHOD_RP_EHRENFEST}[type(method)]
        E.set_problem(code, sim.model.nstates, sim.ndofs, sim.natoms, sim.atoms.masses, dt,
                      algorithm=L.ALG_BCB if code == L.METHOD_RPMD else L.ALG_BCB_TSIT5, nbeads=sim.nbeads,
                      temperature=sim.temperature, **common)
    E.set_model(0, sim.model.kind, sim.model.params())
    if sim.model.kind == L.MODEL_MLP_SURROGATE:
        for k, p in enumerate(sim.model.members):
            E.surrogate_set_params(k, p, sim.model.norm)


def _electronic(dist):
    if isinstance(dist, ProductDistribution):
        return dist.electronic
    if isinstance(dist, dict):
        return dist.get("electronic", PureState(1, Adiabatic()))
    return PureState(1, Adiabatic())


def _configurations(nuc, trajectories, selection):
    """(r, v) stacks (ntraj, flattened column-major) from one DynamicsVariables or a list distribution."""
    if isinstance(nuc, dict):
        r = np.tile(nuc["r"].reshape(-1, order="F"), (trajectories, 1))
        v = np.tile(nuc["v"].reshape(-1, order="F"), (trajectories, 1))
        return r, v
    if isinstance(nuc, DynamicalDistribution) and nuc.is_list():
        idx = np.arange(trajectories) if selection is None else np.asarray(selection) - 1
        r = np.stack([np.asarray(nuc.position[i]).reshape(-1, order="F") for i in idx])
        v = np.stack([np.asarray(nuc.velocity[i]).reshape(-1, order="F") for i in idx])
        return r, v
    return None


def _fermi_dirac(sim, dist):
    """The distribution's FermiDiracState (None without one), checked as DynamicsVariables checks it:
    IESH / EhrenfestNA only, Adiabatic basis, the model's Fermi level (iesh.jl:99-112)."""
    el = _electronic(dist) if isinstance(dist, (ProductDistribution, dict)) else None
    if not isinstance(el, FermiDiracState):
        return None
    if _kind(sim) != "ie":
        raise ValueError("FermiDiracState initial conditions are for IESH / EhrenfestNA simulations")
    if el.diabatic and not isinstance(sim.method, AdiabaticIESH):
        raise ValueError("FermiDiracState in the Diabatic basis is defined for AdiabaticIESH (iesh.jl:138)")
    ef = sim.model.fermi_level
    if abs(ef - el.fermi_level) > np.sqrt(np.finfo(float).eps) * max(abs(ef), abs(el.fermi_level)):
        raise ValueError(f"Fermi level of model and distribution do not match: distribution {el.fermi_level}, "
                         f"model {ef}")
    if el.available_states is not None:
        n = sim.model.nstates
        if max(el.available_states) > n or min(el.available_states) < 1:
            raise ValueError(f"There are only {n} states in the system.")  # get_available_states, DynamicsUtils.jl:211-214
    return el


def _initial_ie(E, sim, dist, trajectories, selection, seed):
    """IESH / EhrenfestNA: the lowest n_electrons adiabatic orbitals occupied, psi[state_e, e] = 1
    (DynamicsVariables(sim, v, r), iesh.jl:89-99, ehrenfest_na.jl:43-48), or with a FermiDiracState
    the thermal occupations drawn on the device (iesh.jl:99-128, ehrenfest_na.jl:43-70)."""
    nuc = dist.nuclear if isinstance(dist, ProductDistribution) else dist
    fd = _fermi_dirac(sim, dist)
    conf = _configurations(nuc, trajectories, selection)
    if conf is not None:
        r, v = conf
        occ = np.tile(np.arange(sim.model.nelectrons, dtype=np.int32), (r.shape[0], 1))
        if sim.nbeads > 1:  # (1, 1, nbeads) per trajectory
            E.load_initial_iesh(r, v, occ)
        else:
            E.load_initial_iesh(r[:, 0], v[:, 0], occ)
    elif sim.nbeads > 1:
        raise ValueError("ring-polymer IESH / EhrenfestNA take bead configurations (DynamicsVariables or a list)")
    elif isinstance(nuc, DynamicalDistribution):
        mr, sr, mv, sv = nuc.moments()
        E.sample_initial(L.SAMPLER_GAUSSIAN, trajectories, [mr[0], sr[0], mv[0], sv[0]], seed)
    else:
        raise ValueError(f"unsupported distribution {type(nuc).__name__}")
    if fd is not None:
        avail = None if fd.available_states is None else np.asarray(fd.available_states) - 1
        E.sample_fermi_dirac(fd.beta, seed, avail, diabatic=fd.diabatic)


def _initial_rp(E, sim, dist, trajectories, selection, seed):
    """Ring polymers: bead configurations (ndofs, natoms, nbeads) uploaded per trajectory."""
    el = _electronic(dist)
    nuc = dist.nuclear if isinstance(dist, ProductDistribution) else dist
    conf = _configurations(nuc, trajectories, selection)
    if conf is None:
        raise ValueError("ring-polymer initial conditions must be DynamicsVariables or a list of bead configurations")
    r, v = conf
    n = max(sim.model.nstates, 2)
    sig = np.zeros((r.shape[0], n, n), complex)
    if sim.model.nstates > 1:
        if el.diabatic:
            raise ValueError("ring-polymer initial conditions need an adiabatic PureState")
        sig[:, el.state - 1, el.state - 1] = 1.0
    state = None if isinstance(sim.method, Classical) else np.full(r.shape[0], el.state - 1, np.int32)
    E.load_initial(r, v, sig, state=state)


def _initial_cme(E, sim, dist, trajectories, selection, seed):
    """CME / BCME: a diabatic PureState (DynamicsVariables(sim, v, r, ::PureState{Diabatic}), cme.jl:20-24)
    and one coordinate, sampled on the device from a Gaussian distribution or uploaded."""
    el = _electronic(dist)
    if not el.diabatic:
        raise ValueError("CME / BCME start from a PureState in the Diabatic basis")
    if el.state not in (1, 2):
        raise ValueError(f"electronic state {el.state} out of range 1..2")
    nuc = dist.nuclear if isinstance(dist, ProductDistribution) else dist
    conf = _configurations(nuc, trajectories, selection)
    if conf is not None:
        r, v = conf
        E.load_initial(r, v, None, state=np.full(r.shape[0], el.state - 1, np.int32))
    elif isinstance(nuc, DynamicalDistribution):
        mr, sr, mv, sv = nuc.moments()
        E.sample_initial(L.SAMPLER_GAUSSIAN, trajectories, [mr[0], sr[0], mv[0], sv[0]], seed,
                         initial_state=el.state - 1)
    else:
        raise ValueError(f"unsupported distribution {type(nuc).__name__}")


def _initial_cl(E, sim, dist, trajectories, selection, seed):
    """Classical dynamics: positions and velocities only (DynamicsVariables(sim, v, r)), sampled on the
    device from a Gaussian distribution or uploaded."""
    nuc = dist.nuclear if isinstance(dist, ProductDistribution) else dist
    conf = _configurations(nuc, trajectories, selection)
    if conf is not None:
        r, v = conf
        E.load_initial(r, v, None)
    elif isinstance(nuc, DynamicalDistribution):
        mr, sr, mv, sv = nuc.moments()
        E.sample_initial(L.SAMPLER_GAUSSIAN, trajectories, [mr[0], sr[0], mv[0], sv[0]], seed)
    else:
        raise ValueError(f"unsupported distribution {type(nuc).__name__}")


def _initial_map(E, sim, dist, trajectories, selection, seed):
    """Spin mapping: a diabatic PureState (DynamicsVariables(sim, v, r, ::PureState{Diabatic}),
    spin_mapping.jl:15-33) and one coordinate, sampled on the device from a Gaussian distribution or
    uploaded; the mapping variables are drawn on the device in the run.  eCMM (cmm.jl:26-34) takes any
    electronic distribution and draws its variables on a sphere."""
    el = _electronic(dist)
    n = sim.model.nstates
    if isinstance(sim.method, eCMM):   # any electronic distribution: eCMM's draw ignores it (cmm.jl:26)
        el = PureState(1)
    if not el.diabatic:
        raise ValueError("spin mapping starts from a PureState in the Diabatic basis")
    if el.state < 1 or el.state > n:
        raise ValueError(f"electronic state {el.state} out of range 1..{n}")
    nuc = dist.nuclear if isinstance(dist, ProductDistribution) else dist
    conf = _configurations(nuc, trajectories, selection)
    if conf is not None:
        r, v = conf
        E.load_initial(r, v, None, state=np.full(r.shape[0], el.state - 1, np.int32))
    elif isinstance(nuc, DynamicalDistribution):
        mr, sr, mv, sv = nuc.moments()
        E.sample_initial(L.SAMPLER_GAUSSIAN, trajectories, [mr[0], sr[0], mv[0], sv[0]], seed,
                         initial_state=el.state - 1, state_basis=1)
    else:
        raise ValueError(f"unsupported distribution {type(nuc).__name__}")


def _initial_rpmap(E, sim, dist, trajectories, selection, seed):
    """NRPMD (a diabatic PureState, nrpmd.jl:49-65) / ring-polymer eCMM (any electronic distribution,
    rpecmm.jl:11-25): bead configurations uploaded; the mapping variables are drawn on the device in the run."""
    el = _electronic(dist)
    n = sim.model.nstates
    if isinstance(sim.method, eCMM):
        el = PureState(1)
    if not el.diabatic:
        raise ValueError("NRPMD starts from a PureState in the Diabatic basis")
    if el.state < 1 or el.state > n:
        raise ValueError(f"electronic state {el.state} out of range 1..{n}")
    nuc = dist.nuclear if isinstance(dist, ProductDistribution) else dist
    conf = _configurations(nuc, trajectories, selection)
    if conf is None:
        raise ValueError("NRPMD / ring-polymer eCMM take bead configurations (DynamicsVariables or a list)")
    r, v = conf
    E.load_initial(r, v, None, state=np.full(r.shape[0], el.state - 1, np.int32))


def _initial_conditions(E, sim, dist, trajectories, selection, seed):
    """Sample on the device where the distribution is recognised, otherwise upload."""
    kind = _kind(sim)
    if kind == "rpmap":
        return _initial_rpmap(E, sim, dist, trajectories, selection, seed)
    if kind == "map":
        return _initial_map(E, sim, dist, trajectories, selection, seed)
    if kind == "ie":
        return _initial_ie(E, sim, dist, trajectories, selection, seed)
    if kind == "rp":
        return _initial_rp(E, sim, dist, trajectories, selection, seed)
    if kind == "cme":
        return _initial_cme(E, sim, dist, trajectories, selection, seed)
    if kind == "cl":
        return _initial_cl(E, sim, dist, trajectories, selection, seed)
    el = _electronic(dist)
    nuc = dist.nuclear if isinstance(dist, ProductDistribution) else dist
    D = sim.ndofs * sim.natoms
    n = sim.model.nstates
    if el.state < 1 or el.state > n:
        raise ValueError(f"electronic state {el.state} out of range 1..{n}")
    state0 = el.state - 1
    if isinstance(nuc, dict):  # a single DynamicsVariables: every trajectory starts there
        r = np.tile(nuc["r"].reshape(-1, order="F"), (trajectories, 1))
        v = np.tile(nuc["v"].reshape(-1, order="F"), (trajectories, 1))
        return _upload(E, sim, r, v, el)
    if isinstance(nuc, DynamicalDistribution) and nuc.is_list():
        idx = np.arange(trajectories) if selection is None else np.asarray(selection) - 1
        r = np.stack([np.asarray(nuc.position[i]).reshape(-1, order="F") for i in idx])
        v = np.stack([np.asarray(nuc.velocity[i]).reshape(-1, order="F") for i in idx])
        return _upload(E, sim, r, v, el)
    if isinstance(nuc, DynamicalDistribution):
        mr, sr, mv, sv = nuc.moments()
        if sim.model.kind == L.MODEL_SPIN_BOSON:
            prm = np.concatenate([mr, sr, mv, sv])
            E.sample_initial(L.SAMPLER_HARMONIC_WIGNER, trajectories, prm, seed, initial_state=state0,
                             state_basis=1 if el.diabatic else 0)
            return
        prm = np.stack([mr, sr, mv, sv], axis=1).reshape(-1)
        E.sample_initial(L.SAMPLER_GAUSSIAN, trajectories, prm, seed, initial_state=state0,
                         state_basis=1 if el.diabatic else 0)
        return
    raise ValueError(f"unsupported distribution {type(nuc).__name__}")


def _upload(E, sim, r, v, el):
    n = sim.model.nstates
    ntraj = r.shape[0]
    sig = np.zeros((ntraj, n, n), complex)
    if el.diabatic:
        raise ValueError("uploaded initial conditions need an adiabatic PureState (or use a distribution)")
    sig[:, el.state - 1, el.state - 1] = 1.0
    E.load_initial(r, v, sig, state=np.full(ntraj, el.state - 1, np.int32))


def _results(E, t0, dt, save_every, nsteps):
    res = {"r": E.download(L.FIELD_R), "v": E.download(L.FIELD_V),
           "sigma": E.download(L.FIELD_SIGMA_RE) + 1j * E.download(L.FIELD_SIGMA_IM),
           "state": E.download(L.FIELD_STATE), "steps": E.download(L.FIELD_STEPS),
           "nhops": E.download(L.FIELD_NHOPS), "pop_ad": E.download(L.FIELD_SAVE_POP_AD),
           "pop_di": E.download(L.FIELD_SAVE_POP_DI), "save_state": E.download(L.FIELD_SAVE_STATE)}
    try:  # energies per save exist for the two-state one-dof problems
        res["ekin"], res["epot"] = E.download(L.FIELD_SAVE_EKIN), E.download(L.FIELD_SAVE_EPOT)
    except L.NQCDError:
        pass
    return res


def save_schedule(saveat, t0, t1, dt):
    """Save slots for SciML's saveat (which run_dynamics.jl:91-97 forwards): None saves every step; a number
    d saves t0:d:t1; a sequence must start at t0 and be evenly spaced by a multiple of dt.  Returns
    (save_every, nsave, times); the device saves slot k at step k * save_every."""
    nsteps = int(round((t1 - t0) / dt))
    tol = 1e-9 * max(1.0, abs(dt))

    def multiple(d):
        k = int(round(d / dt))
        if k < 1 or abs(k * dt - d) > tol * k:
            raise ValueError(f"saveat spacing {d} is not a positive multiple of dt = {dt}")
        return k

    if saveat is None:
        k, nsave = 1, nsteps + 1
    elif np.isscalar(saveat):
        k = multiple(float(saveat))
        if nsteps % k:
            raise ValueError("t1 - t0 must be a multiple of saveat")
        nsave = nsteps // k + 1
    else:
        ts = np.asarray(saveat, dtype=np.float64).ravel()
        if ts.size == 0:
            raise ValueError("empty saveat")
        if abs(ts[0] - t0) > tol:
            raise ValueError("saveat must start at tspan[0]")
        if ts.size == 1:
            return 0, 1, ts
        d = np.diff(ts)
        if np.any(np.abs(d - d[0]) > tol * ts.size):
            raise ValueError("saveat times must be evenly spaced (non-uniform saveat is not supported)")
        k = multiple(float(d[0]))
        if (ts.size - 1) * k > nsteps:
            raise ValueError("saveat extends past tspan[1]")
        return k, int(ts.size), ts
    return k, nsave, t0 + dt * k * np.arange(nsave)


def run_dynamics(sim, tspan, distribution, *, output, selection=None, reduction=None, ensemble_algorithm=None,
                 trajectories=1, savetime=True, precompile_dynamics=False, dt=1.0, saveat=None, callback=None,
                 **kwargs):
    """Run `trajectories` trajectories over `tspan` sampling `distribution` (run_dynamics.jl:42-135)."""
    if kwargs:
        raise TypeError(f"unsupported keyword arguments for EnsembleMI355X: {sorted(kwargs)}")
    if ensemble_algorithm is None:
        ensemble_algorithm = EnsembleMI355X()
    if not isinstance(ensemble_algorithm, EnsembleMI355X):
        raise ValueError("this package provides only the EnsembleMI355X ensemble algorithm")
    if reduction is None:
        reduction = SortByTrajectoryReduction()
    outputs = output if isinstance(output, tuple) else (output,)
    t0, t1 = float(tspan[0]), float(tspan[1])
    nsteps = int(round((t1 - t0) / dt))
    save_every, nsave_user, times = save_schedule(saveat, t0, t1, dt)
    term, cell = _callbacks(callback)
    kind = _kind(sim)
    if kind not in ("sh", "cme", "cl") and callback is not None:
        raise ValueError("termination and cell callbacks are supported for FSSH / Ehrenfest / CME / BCME / classical "
                         "simulations")
    if not isinstance(getattr(sim, "cell", None), PeriodicCell):
        cell = L.CELL_NONE   # InfiniteCell: every atom is inside, the cell callbacks never fire
    if int(trajectories) < 1:
        raise ValueError("trajectories must be >= 1")
    _fermi_dirac(sim, distribution)
    if kind == "ie" and sim.nbeads > 1:
        nuc = distribution.nuclear if isinstance(distribution, ProductDistribution) else distribution
        if not (isinstance(nuc, dict) or (isinstance(nuc, DynamicalDistribution) and nuc.is_list())):
            raise ValueError("ring-polymer IESH / EhrenfestNA take bead configurations (DynamicsVariables or a list)")
    log.info("Performing %d trajectories on MI355X device %d.", trajectories, ensemble_algorithm.device)
    E = L.Ensemble(ensemble_algorithm.device)
    E._schedule = ensemble_algorithm.schedule
    try:
        _problem(E, sim, dt)
        if cell != L.CELL_NONE:
            E.set_cell(cell, sim.cell.vectors, sim.cell.periodicity)
        _initial_conditions(E, sim, distribution, int(trajectories), selection, ensemble_algorithm.seed)
        if getattr(ensemble_algorithm, "autotune", False) and kind == "sh":
            try:
                ensemble_algorithm.tuning = E.autotune(t0, t0 + dt * min(nsteps, ensemble_algorithm.autotune_steps),
                                                       ensemble_algorithm.seed, term=term)
            except L.NQCDError as e:
                if e.code != L.ERR_UNSUPPORTED:
                    raise
        stats = E.run(t0, t1, ensemble_algorithm.seed, save_every=save_every, term=term)
        result = _reduce(E, sim, outputs, reduction, times, savetime, t0, dt, save_every, nsteps, nsave_user)
    finally:
        E.close()
    if isinstance(reduction, SortByTrajectoryReduction) and trajectories == 1:
        return result[0]
    return result


def _per_traj_results(E, sim, t0, dt, save_every, nsteps, nsave=None):
    """Per-trajectory dictionaries of the device fields (per-save arrays cut to the first nsave slots)."""
    out = _per_traj_fields(E, sim, t0, dt, save_every, nsteps)
    if nsave is not None and nsave < E.nsave:
        for d in out:
            for k in ("pop_ad", "pop_di", "save_state", "centroid", "ekin", "epot", "save_r", "save_v"):
                if k in d:
                    d[k] = d[k][:nsave]
    return out


def _per_traj_fields(E, sim, t0, dt, save_every, nsteps):
    kind = _kind(sim)
    if kind == "ie":
        res = {"r": E.download(L.FIELD_R), "v": E.download(L.FIELD_V),
               "psi": E.download(L.FIELD_SIGMA_RE) + 1j * E.download(L.FIELD_SIGMA_IM),
               "state": E.download(L.FIELD_STATE), "steps": E.download(L.FIELD_STEPS),
               "nhops": E.download(L.FIELD_NHOPS), "pop_ad": E.download(L.FIELD_SAVE_POP_AD),
               "pop_di": E.download(L.FIELD_SAVE_POP_DI)}
        return [{"r": res["r"][i].reshape(sim.size()), "v": res["v"][i].reshape(sim.size()), "sigma": res["psi"][i],
                 "state": res["state"][i], "steps": int(res["steps"][i]), "nhops": int(res["nhops"][i]),
                 "pop_ad": res["pop_ad"][i], "pop_di": res["pop_di"][i], "final_pop_ad": res["pop_ad"][i][-1],
                 "t0": t0, "dt": dt} for i in range(E.ntraj)]
    if kind == "cl":
        r, v, stp = E.download(L.FIELD_R), E.download(L.FIELD_V), E.download(L.FIELD_STEPS)
        sr, sv = E.download(L.FIELD_SAVE_R), E.download(L.FIELD_SAVE_V)
        ek, ep = E.download(L.FIELD_SAVE_EKIN), E.download(L.FIELD_SAVE_EPOT)
        shp = sim.size()
        return [{"r": r[i].reshape(shp, order="F"), "v": v[i].reshape(shp, order="F"), "steps": int(stp[i]),
                 "save_r": [x.reshape(shp, order="F") for x in sr[i]],
                 "save_v": [x.reshape(shp, order="F") for x in sv[i]], "ekin": ek[i], "epot": ep[i],
                 "t0": t0, "dt": dt} for i in range(E.ntraj)]
    if kind == "rpmap":   # qmap / pmap (n, nbeads) per trajectory, as the reference's ComponentVector holds them
        r, v, X, P = E.download(L.FIELD_R), E.download(L.FIELD_V), E.download(L.FIELD_MAP_X), E.download(L.FIELD_MAP_P)
        pd, ek, ep = E.download(L.FIELD_SAVE_POP_DI), E.download(L.FIELD_SAVE_EKIN), E.download(L.FIELD_SAVE_EPOT)
        cen, k0 = E.download(L.FIELD_SAVE_CENTROID), E.download(L.FIELD_INIT_POP_DI)
        norm = float(sim.model.nstates) if isinstance(sim.method, eCMM) else 1.0
        shp = sim.size()
        return [{"r": r[i].reshape(shp, order="F"), "v": v[i].reshape(shp, order="F"), "qmap": X[i].T.copy(),
                 "pmap": P[i].T.copy(), "pop_di": pd[i], "ekin": ek[i], "epot": ep[i], "centroid": cen[i],
                 "pop_init": k0[i], "corr_norm": norm, "steps": nsteps, "t0": t0, "dt": dt} for i in range(E.ntraj)]
    if kind == "map":
        r, v, X, P = E.download(L.FIELD_R), E.download(L.FIELD_V), E.download(L.FIELD_MAP_X), E.download(L.FIELD_MAP_P)
        pd, ek, ep = E.download(L.FIELD_SAVE_POP_DI), E.download(L.FIELD_SAVE_EKIN), E.download(L.FIELD_SAVE_EPOT)
        k0 = E.download(L.FIELD_INIT_POP_DI)
        norm = float(sim.model.nstates) if isinstance(sim.method, eCMM) else 1.0
        shp = sim.size()
        return [{"r": r[i].reshape(shp, order="F"), "v": v[i].reshape(shp, order="F"), "qmap": X[i], "pmap": P[i],
                 "pop_di": pd[i], "ekin": ek[i], "epot": ep[i], "pop_init": k0[i], "corr_norm": norm,
                 "steps": nsteps, "t0": t0, "dt": dt} for i in range(E.ntraj)]
    if kind == "cme":
        r, v, st = E.download(L.FIELD_R), E.download(L.FIELD_V), E.download(L.FIELD_STATE)
        nh, stp = E.download(L.FIELD_NHOPS), E.download(L.FIELD_STEPS)
        sr, ek, ep = E.download(L.FIELD_SAVE_R), E.download(L.FIELD_SAVE_EKIN), E.download(L.FIELD_SAVE_EPOT)
        ss = E.download(L.FIELD_SAVE_STATE)
        shp = sim.size()
        return [{"r": r[i].reshape(shp, order="F"), "v": v[i].reshape(shp, order="F"), "state": int(st[i]),
                 "steps": int(stp[i]), "nhops": int(nh[i]), "save_r": [x.reshape(shp, order="F") for x in sr[i]],
                 "ekin": ek[i], "epot": ep[i], "save_state": ss[i], "sigma": None,
                 "final_pop_ad": np.eye(2)[st[i]], "final_pop_di": np.eye(2)[st[i]], "t0": t0, "dt": dt}
                for i in range(E.ntraj)]
    if kind == "rp":
        r, v = E.download(L.FIELD_R), E.download(L.FIELD_V)
        cen = E.download(L.FIELD_SAVE_CENTROID)
        out = []
        elec = sim.model.nstates > 1
        if elec:
            sig = E.download(L.FIELD_SIGMA_RE) + 1j * E.download(L.FIELD_SIGMA_IM)
            st, nh = E.download(L.FIELD_STATE), E.download(L.FIELD_NHOPS)
            pa, pd, ss = E.download(L.FIELD_SAVE_POP_AD), E.download(L.FIELD_SAVE_POP_DI), E.download(L.FIELD_SAVE_STATE)
        for i in range(E.ntraj):
            d = {"r": r[i].reshape(sim.size(), order="F"), "v": v[i].reshape(sim.size(), order="F"),
                 "centroid": cen[i], "steps": nsteps, "t0": t0, "dt": dt}
            if elec:
                fin = np.zeros(2)
                if isinstance(sim.method, FSSH):
                    fin[st[i]] = 1.0
                else:
                    fin = np.real(np.diag(sig[i]))
                d.update({"sigma": sig[i], "state": int(st[i]), "nhops": int(nh[i]), "pop_ad": pa[i], "pop_di": pd[i],
                          "save_state": ss[i], "final_pop_ad": fin})
            out.append(d)
        return out
    res = _results(E, t0, dt, save_every, nsteps)
    n = sim.model.nstates
    fin = np.zeros((E.ntraj, n))
    if isinstance(sim.method, FSSH):
        fin[np.arange(E.ntraj), res["state"]] = 1.0
    else:
        fin = np.real(np.einsum("tii->ti", res["sigma"]))
    res["final_pop_ad"] = fin
    try:  # diag(U sigma' U^T) of the final state (fssh.jl:132-142, ehrenfest.jl:75-83): two-state one-dof problems
        fin_di = E.download(L.FIELD_FINAL_POP_DI)
    except L.NQCDError:
        fin_di = None
    out = []
    for i in range(E.ntraj):
        out.append({"r": res["r"][i].reshape(sim.size(), order="F"), "v": res["v"][i].reshape(sim.size(), order="F"),
                    "sigma": res["sigma"][i], "state": int(res["state"][i]), "steps": int(res["steps"][i]),
                    "nhops": int(res["nhops"][i]), "pop_ad": res["pop_ad"][i], "pop_di": res["pop_di"][i],
                    "save_state": res["save_state"][i], "final_pop_ad": fin[i], "t0": t0, "dt": dt})
        if fin_di is not None:
            out[-1]["final_pop_di"] = fin_di[i]
        if "ekin" in res:
            out[-1].update({"ekin": res["ekin"][i], "epot": res["epot"][i]})
    return out


def _reduce(E, sim, outputs, reduction, times, savetime, t0, dt, save_every, nsteps, nsave):
    """The reduction's result over the device results; per-save outputs keep the first nsave slots (a
    saveat sequence that stops before t1)."""
    n = sim.model.nstates
    if isinstance(reduction, (MeanReduction, SumReduction)):
        red = L.RED_MEAN if isinstance(reduction, MeanReduction) else L.RED_SUM
        result = {"Time": times} if savetime else {}
        host = None
        for f in outputs:
            kind = f.device_reduction if isinstance(f, Output) else None
            if kind is not None:
                if kind in (L.OUT_ADIABATIC_POPULATION, L.OUT_DIABATIC_POPULATION):
                    v = E.reduce(kind, red, n * E.nsave).reshape(E.nsave, n)
                    result[output_name(f)] = list(v[:nsave])
                elif kind in (L.OUT_SCATTERING_1D, L.OUT_SCATTERING_1D_DIABATIC):
                    v = E.reduce(kind, red, 2 * n)
                    result[output_name(f)] = {"reflection": v[:n], "transmission": v[n:]}
                elif kind in (L.OUT_POSITION, L.OUT_VELOCITY):
                    v = E.reduce(kind, red, E.nsave * E.D)
                    result[output_name(f)] = [x.reshape(sim.size(), order="F") for x in v.reshape(E.nsave, E.D)[:nsave]]
                elif kind in (L.OUT_CENTROID_POSITION, L.OUT_KINETIC_ENERGY, L.OUT_TOTAL_ENERGY):
                    result[output_name(f)] = list(E.reduce(kind, red, E.nsave)[:nsave])
                elif kind == L.OUT_POPULATION_CORRELATION:
                    v = E.reduce(kind, red, n * n * E.nsave).reshape(E.nsave, n, n)
                    result[output_name(f)] = [m.T for m in v[:nsave]]  # column-major n x n per save
                continue
            if host is None:
                host = _per_traj_results(E, sim, t0, dt, save_every, nsteps, nsave)
            vals = [f.per_trajectory(h) for h in host]
            tot = _tree_sum(vals)
            result[output_name(f)] = _scale(tot, 1.0 / len(vals)) if red == L.RED_MEAN else tot
        return result
    host = _per_traj_results(E, sim, t0, dt, save_every, nsteps, nsave)
    per = []
    for h in host:
        d = {"Time": times} if savetime else {}
        for f in outputs:
            d[output_name(f)] = f.per_trajectory(h) if isinstance(f, Output) else f(h, None)
        per.append(d)
    if isinstance(reduction, FileReduction):  # reductions.jl:72-91, one batch of every trajectory
        return reduction([], per, range(1, len(per) + 1))[0]
    if isinstance(reduction, SortByOutputReduction):
        keys = per[0].keys() if per else []
        return {k: [d[k] for d in per] for k in keys}
    return per


def _tree_sum(vals):
    first = vals[0]
    if isinstance(first, dict):
        return {k: _tree_sum([v[k] for v in vals]) for k in first}
    if isinstance(first, list):
        return [_tree_sum([v[j] for v in vals]) for j in range(len(first))]
    return np.sum(np.asarray(vals, dtype=np.float64), axis=0)


def _scale(x, s):
    if isinstance(x, dict):
        return {k: _scale(v, s) for k, v in x.items()}
    if isinstance(x, list):
        return [_scale(v, s) for v in x]
    return x * s
