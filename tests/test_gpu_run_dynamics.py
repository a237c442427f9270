"""run_dynamics (the mirrored reference API, run_dynamics.jl:42-135) for FSSH and Ehrenfest on the
Tully models: every output (DynamicsOutputs.jl:178-243,313-338, TimeCorrelationFunctions.jl:26-88)
under every reduction (reductions.jl:8-297) against the FP64 oracle on the same initial conditions
and hop uniforms.  97 trajectories (not a multiple of the wavefront), termination at |x| > 10."""
import numpy as np
import pytest

import nqcd_mi355x as N

pytestmark = pytest.mark.gpu

TULLY = {0: [0.01, 1.6, 0.005, 1.0], 1: [0.1, 0.28, 0.015, 0.06, 0.05], 2: [6e-4, 0.1, 0.9]}
MODELS = {0: N.TullyModelOne, 1: N.TullyModelTwo, 2: N.TullyModelThree}
MASS = 2000.0
NTRAJ, NSTEPS, EVERY = 97, 3000, 100


def _case(oracle, key, method_cls, kind, outputs, reduction, seed=3, ens=None):
    rng = np.random.default_rng(seed)
    p0 = rng.uniform(8.0, 30.0, size=NTRAJ)
    r0 = np.full((NTRAJ, 1), -8.0)
    v0 = (p0 / MASS)[:, None]
    sim = N.Simulation[method_cls](N.Atoms(MASS), MODELS[kind](*TULLY[kind]))
    dist = N.DynamicalDistribution([np.array([[v]]) for v in v0[:, 0]], [np.array([[-8.0]])] * NTRAJ, (1, 1))
    res = N.run_dynamics(sim, (0.0, float(NSTEPS)), dist * N.PureState(1, N.Adiabatic()), trajectories=NTRAJ,
                         dt=1.0, saveat=float(EVERY), output=outputs, reduction=reduction,
                         callback=N.TerminatingCallback(1, 10.0),
                         ensemble_algorithm=ens or N.EnsembleMI355X(seed=key))
    m = oracle.make_model(kind, 2, 1, TULLY[kind])
    sig0 = np.zeros((NTRAJ, 2, 2), complex); sig0[:, 0, 0] = 1.0
    method = 0 if method_cls is N.FSSH else 1
    c = oracle.run_ensemble(method, m, MASS, r0, v0, sig0, state0=np.zeros(NTRAJ, np.int32), dt=1.0, nsteps=NSTEPS,
                            key=key, save_every=EVERY, nsave=NSTEPS // EVERY + 1, term_x=10.0, energies=True)
    return sim, res, c


def test_fssh_mean_reduction_all_outputs(oracle, key):
    outs = (N.OutputAdiabaticPopulation, N.OutputDiabaticPopulation, N.OutputSurfaceHops, N.OutputFinalTime)
    sim, res, c = _case(oracle, key, N.FSSH, 0, outs, N.MeanReduction())
    assert c["nhops"].sum() > 0
    nsave = NSTEPS // EVERY + 1
    assert np.allclose(res["Time"], np.arange(nsave) * EVERY)
    assert np.allclose(np.array(res["OutputAdiabaticPopulation"]), c["pop_ad"].mean(axis=0), rtol=1e-9, atol=1e-12)
    assert np.allclose(np.array(res["OutputDiabaticPopulation"]), c["pop_di"].mean(axis=0), rtol=1e-9, atol=1e-12)
    assert np.isclose(res["OutputSurfaceHops"], c["nhops"].mean(), rtol=1e-12)
    assert np.isclose(res["OutputFinalTime"], (c["steps"] * 1.0).mean(), rtol=1e-12)


def test_fssh_scattering_and_correlation_sum_reduction(oracle, key):
    """OutputStateResolvedScattering1D(:adiabatic) splits the final one-hot population by x > 0;
    PopulationCorrelationFunction is P_diabatic(0) P_diabatic(t)^T per save; Sum reduction."""
    sim = N.Simulation[N.FSSH](N.Atoms(MASS), N.TullyModelOne(*TULLY[0]))
    outs = (N.OutputStateResolvedScattering1D(sim, "adiabatic"), N.PopulationCorrelationFunction(sim))
    sim, res, c = _case(oracle, key, N.FSSH, 0, outs, N.SumReduction())
    fin = np.zeros((NTRAJ, 2)); fin[np.arange(NTRAJ), c["state"]] = 1.0
    right = c["r"][:, 0] > 0
    sc = res["OutputStateResolvedScattering1D"]
    assert np.allclose(sc["transmission"], fin[right].sum(axis=0), atol=1e-12)
    assert np.allclose(sc["reflection"], fin[~right].sum(axis=0), atol=1e-12)
    assert np.isclose(np.sum(sc["transmission"]) + np.sum(sc["reflection"]), NTRAJ)
    pcf = np.array(res["PopulationCorrelationFunction"])
    ref = np.einsum("ta,tsb->sab", c["pop_di"][:, 0, :], c["pop_di"])
    assert pcf.shape == ref.shape
    assert np.allclose(pcf, ref, rtol=1e-9, atol=1e-11)


def test_fssh_per_trajectory_reductions(oracle, key):
    """SortByTrajectoryReduction gives one dictionary per trajectory and SortByOutputReduction one
    list per output, both equal to the oracle trajectory by trajectory."""
    outs = (N.OutputDiscreteState, N.OutputFinal)
    _, per, c = _case(oracle, key, N.FSSH, 2, outs, N.SortByTrajectoryReduction(), seed=4)
    assert len(per) == NTRAJ
    for i in (0, 1, NTRAJ // 2, NTRAJ - 1):
        assert np.array_equal(np.array(per[i]["OutputDiscreteState"]), c["save_state"][i] + 1)
        fin = per[i]["OutputFinal"]
        assert fin["state"] == c["state"][i] + 1
        assert np.allclose(fin["r"].ravel(), c["r"][i], rtol=1e-9, atol=1e-12)
        assert np.allclose(fin["v"].ravel(), c["v"][i], rtol=1e-9, atol=1e-14)
        assert np.allclose(fin["sigma"], c["sig_re"][i] + 1j * c["sig_im"][i], rtol=1e-9, atol=1e-11)
    _, byout, _ = _case(oracle, key, N.FSSH, 2, outs, N.SortByOutputReduction(), seed=4)
    assert set(byout) == {"Time", "OutputDiscreteState", "OutputFinal"}
    assert len(byout["OutputDiscreteState"]) == NTRAJ
    states = np.array([np.array(s) for s in byout["OutputDiscreteState"]])
    assert np.array_equal(states, c["save_state"] + 1)


def test_ehrenfest_populations_mean_reduction(oracle, key):
    outs = (N.OutputAdiabaticPopulation, N.OutputDiabaticPopulation)
    _, res, c = _case(oracle, key, N.Ehrenfest, 1, outs, N.MeanReduction(), seed=5)
    assert np.allclose(np.array(res["OutputAdiabaticPopulation"]), c["pop_ad"].mean(axis=0), rtol=1e-9, atol=1e-12)
    assert np.allclose(np.array(res["OutputDiabaticPopulation"]), c["pop_di"].mean(axis=0), rtol=1e-9, atol=1e-12)
    # mean-field populations are not one-hot and stay normalised; the save slots of a trajectory
    # after its termination read zero, so each save's total is the fraction still running
    pa = np.array(res["OutputAdiabaticPopulation"])
    alive = np.array([np.mean(c["steps"] >= s * EVERY) for s in range(len(pa))])
    assert alive[-1] < 1.0
    assert np.allclose(pa.sum(axis=1), alive, atol=1e-12)
    assert np.any((pa[5] > 1e-3) & (pa[5] < 1 - 1e-3))


@pytest.mark.parametrize("method_cls,kind", [(N.FSSH, 0), (N.Ehrenfest, 1)])
def test_energy_outputs_mean_and_per_trajectory(oracle, key, method_cls, kind):
    """OutputKineticEnergy and OutputTotalEnergy (DynamicsOutputs.jl:90-98) at every save: the device
    Mean reduction and the per-trajectory lists against the oracle's saved energies."""
    outs = (N.OutputKineticEnergy, N.OutputTotalEnergy)
    _, res, c = _case(oracle, key, method_cls, kind, outs, N.MeanReduction(), seed=6)
    assert np.allclose(np.array(res["OutputKineticEnergy"]), c["ekin"].mean(axis=0), rtol=1e-9, atol=1e-14)
    assert np.allclose(np.array(res["OutputTotalEnergy"]), (c["ekin"] + c["epot"]).mean(axis=0), rtol=1e-9,
                       atol=1e-14)
    _, per, _ = _case(oracle, key, method_cls, kind, outs, N.SortByTrajectoryReduction(), seed=6)
    for i in (0, NTRAJ // 3, NTRAJ - 1):
        n = c["steps"][i] // EVERY + 1  # save slots the trajectory reached before it terminated
        assert np.allclose(np.array(per[i]["OutputKineticEnergy"])[:n], c["ekin"][i, :n], rtol=1e-9, atol=1e-14)
        assert np.allclose(np.array(per[i]["OutputTotalEnergy"])[:n], (c["ekin"] + c["epot"])[i, :n], rtol=1e-9,
                           atol=1e-14)


def test_autotuned_run_dynamics_matches_oracle(oracle, key):
    """EnsembleMI355X(autotune=True) tunes the persistent kernel's launch shape on the first steps and
    then runs; the outputs are those of the oracle as for the default shape."""
    ens = N.EnsembleMI355X(seed=key, autotune=True, autotune_steps=100)
    outs = (N.OutputAdiabaticPopulation, N.OutputSurfaceHops)
    _, res, c = _case(oracle, key, N.FSSH, 1, outs, N.MeanReduction(), seed=7, ens=ens)
    assert ens.tuning is not None and ens.tuning["family"] == 1 and ens.tuning["n_candidates"] >= 6
    assert np.allclose(np.array(res["OutputAdiabaticPopulation"]), c["pop_ad"].mean(axis=0), rtol=1e-9, atol=1e-12)
    assert np.isclose(res["OutputSurfaceHops"], c["nhops"].mean(), rtol=1e-12)


def test_single_trajectory_returns_its_dictionary(oracle, key):
    """trajectories = 1 with the default reduction returns the one trajectory's dictionary."""
    sim = N.Simulation[N.FSSH](N.Atoms(MASS), N.TullyModelOne(*TULLY[0]))
    dist = N.DynamicalDistribution(20.0 / MASS, -8.0, (1, 1)) * N.PureState(1, N.Adiabatic())
    out = N.run_dynamics(sim, (0.0, 2000.0), dist, trajectories=1, dt=1.0, output=N.OutputFinal,
                         callback=N.TerminatingCallback(1, 10.0), ensemble_algorithm=N.EnsembleMI355X(seed=key))
    assert isinstance(out, dict) and "OutputFinal" in out and "Time" in out
    m = oracle.make_model(0, 2, 1, TULLY[0])
    sig0 = np.zeros((1, 2, 2), complex); sig0[:, 0, 0] = 1.0
    c = oracle.run_ensemble(0, m, MASS, np.array([[-8.0]]), np.array([[20.0 / MASS]]), sig0,
                            state0=np.zeros(1, np.int32), dt=1.0, nsteps=2000, key=key, term_x=10.0)
    assert np.allclose(out["OutputFinal"]["r"].ravel(), c["r"][0], rtol=1e-9)
    assert out["OutputFinal"]["state"] == c["state"][0] + 1


@pytest.mark.parametrize("method_cls", [N.FSSH, N.Ehrenfest])
def test_diabatic_scattering_at_the_final_state(oracle, key, method_cls):
    """OutputStateResolvedScattering1D(sim, :diabatic) (DynamicsOutputs.jl:313-338) takes the diabatic
    populations of each trajectory's final state, also for trajectories that terminated between save slots:
    the device reduction and the per-trajectory path against the oracle's diabatic population at the step
    each trajectory ended on (saved at every step)."""
    sim = N.Simulation[method_cls](N.Atoms(MASS), N.TullyModelOne(*TULLY[0]))
    out = N.OutputStateResolvedScattering1D(sim, "diabatic")
    _, res_sum, _ = _case(oracle, key, method_cls, 0, (out,), N.SumReduction())
    _, res_traj, _ = _case(oracle, key, method_cls, 0, (out,), N.SortByTrajectoryReduction())
    rng = np.random.default_rng(3)
    p0 = rng.uniform(8.0, 30.0, size=NTRAJ)
    r0 = np.full((NTRAJ, 1), -8.0); v0 = (p0 / MASS)[:, None]
    sig0 = np.zeros((NTRAJ, 2, 2), complex); sig0[:, 0, 0] = 1.0
    m = oracle.make_model(0, 2, 1, TULLY[0])
    c = oracle.run_ensemble(0 if method_cls is N.FSSH else 1, m, MASS, r0, v0, sig0, state0=np.zeros(NTRAJ, np.int32),
                            dt=1.0, nsteps=NSTEPS, key=key, save_every=1, nsave=NSTEPS + 1, term_x=10.0)
    fin = c["pop_di"][np.arange(NTRAJ), c["steps"]]
    right = c["r"][:, 0] > 0
    assert np.allclose(res_sum["OutputStateResolvedScattering1D"]["transmission"], fin[right].sum(axis=0), rtol=1e-9)
    assert np.allclose(res_sum["OutputStateResolvedScattering1D"]["reflection"], fin[~right].sum(axis=0), rtol=1e-9)
    for i, d in enumerate(res_traj):
        sc = d["OutputStateResolvedScattering1D"]
        assert np.allclose(sc["transmission" if right[i] else "reflection"], fin[i], rtol=1e-9, atol=1e-12)
    assert np.isclose(sum(np.sum(d["OutputStateResolvedScattering1D"]["transmission"]) +
                          np.sum(d["OutputStateResolvedScattering1D"]["reflection"]) for d in res_traj), NTRAJ)


def test_saveat_every_step_and_sequences(oracle, key):
    """saveat forms (run_dynamics.jl:91-97 forwards them to SciML's solve): none saves every step, a
    uniform sequence from t0 keeps its own times and slots, and the per-save outputs equal the oracle's."""
    sim = N.Simulation[N.FSSH](N.Atoms(MASS), N.TullyModelOne(*TULLY[0]))
    dist = N.DynamicalDistribution(20.0 / MASS, -8.0, (1, 1)) * N.PureState(1, N.Adiabatic())
    nsteps = 400
    full = N.run_dynamics(sim, (0.0, float(nsteps)), dist, trajectories=3, dt=1.0, output=N.OutputAdiabaticPopulation,
                          reduction=N.MeanReduction(), ensemble_algorithm=N.EnsembleMI355X(seed=key))
    assert len(full["Time"]) == nsteps + 1 and len(full["OutputAdiabaticPopulation"]) == nsteps + 1
    seq = N.run_dynamics(sim, (0.0, float(nsteps)), dist, trajectories=3, dt=1.0, saveat=np.arange(0.0, 151.0, 10.0),
                         output=N.OutputAdiabaticPopulation, reduction=N.MeanReduction(),
                         ensemble_algorithm=N.EnsembleMI355X(seed=key))
    assert np.allclose(seq["Time"], np.arange(0.0, 151.0, 10.0))
    assert np.allclose(np.array(seq["OutputAdiabaticPopulation"]), np.array(full["OutputAdiabaticPopulation"])[::10][:16])
