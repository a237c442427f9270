"""Spin-mapping dynamics integrated by MInt on the device (Simulation{SpinMappingW}, spin_mapping.jl; mint.jl)
against the FP64 oracle: mapping variables drawn on the device identical to the oracle's draws, trajectories
(positions, velocities, mapping variables, per-save diabatic populations and energies) to 1e-9 over a few
hundred steps on two-, three- and four-state models, the device reductions against the per-trajectory
fields, and at 1e5 trajectories the properties that do not depend on the ensemble size: the total
population sum_i (X_i^2 + P_i^2 - gamma) / 2 = 1 kept by the orthogonal MInt rotation, and the energy of
every trajectory conserved to the integrator's accuracy."""
import numpy as np
import pytest

from nstate_models import morse_table

pytestmark = pytest.mark.gpu

TULLY1 = [0.01, 1.6, 0.005, 1.0]
DW = [1 / 1200, 1 / 40, 0.01, 0.005, 1.0]


def _case(L, O, name):
    """(device model kind, params, oracle model, n, mass, x0 mean, x0 sd, v0 mean, v0 sd, dt)"""
    if name == "tully1":
        return L.MODEL_TULLY1, TULLY1, O.make_model(O.MODEL_TULLY1, 2, 1, TULLY1), 2, 2000.0, -4.0, 0.5, 0.01, 1e-3, 1.0
    if name == "double_well":
        return L.MODEL_DW_GAUSSIAN, DW, O.make_model(O.MODEL_DW_GAUSSIAN, 2, 1, DW), 2, 1836.0, -4.5, 0.5, 0.0, 1e-3, 2.0
    n = 3 if name == "morse3" else 4
    tab = list(morse_table(n))
    return L.MODEL_MORSE_N, [n] + tab, O.make_model(O.MODEL_MORSE_N, n, 1, tab), n, 20000.0, 4.1, 0.1, 0.0, 2e-4, 2.0


def _device(L, kind, params, n, mass, dt, r0, v0, state, nsteps, save_every, key, X=None, P=None, ecmm_gamma=None):
    E = L.Ensemble(0)
    if ecmm_gamma is None:
        E.set_problem(L.METHOD_SPIN_MAPPING, n, 1, 1, [mass], dt, algorithm=L.ALG_MINT)
    else:
        E.set_problem(L.METHOD_ECMM, n, 1, 1, [mass], dt, algorithm=L.ALG_MINT, mapping_gamma=ecmm_gamma)
    E.set_model(0, kind, params)
    E.load_initial(r0.reshape(-1, 1), v0.reshape(-1, 1), None, state=state)
    if X is not None:
        E.load_mapping(X, P)
    E.run(0.0, dt * nsteps, key, save_every=save_every)
    g = {f: E.download(getattr(L, "FIELD_" + f)) for f in ("R", "V", "MAP_X", "MAP_P", "SAVE_POP_DI", "SAVE_EKIN",
                                                             "SAVE_EPOT")}
    g["red_pop"] = E.reduce(L.OUT_DIABATIC_POPULATION, L.RED_MEAN, n * E.nsave).reshape(E.nsave, n)
    g["red_ekin"] = E.reduce(L.OUT_KINETIC_ENERGY, L.RED_SUM, E.nsave)
    g["red_etot"] = E.reduce(L.OUT_TOTAL_ENERGY, L.RED_MEAN, E.nsave)
    g["noalloc_pop"] = E.reduce_noalloc(L.OUT_DIABATIC_POPULATION, L.RED_MEAN, n * E.nsave).reshape(E.nsave, n)
    # column-major n x n per slot: element [slot, b, a] = norm K_a(0) P_b(t_slot)
    g["red_corr"] = E.reduce(L.OUT_POPULATION_CORRELATION, L.RED_MEAN, n * n * E.nsave).reshape(E.nsave, n, n)
    g["noalloc_corr"] = E.reduce_noalloc(L.OUT_POPULATION_CORRELATION, L.RED_MEAN, n * n * E.nsave).reshape(E.nsave, n, n)
    g["stats"] = E.stats()
    E.close()
    return g


@pytest.mark.parametrize("name", ["tully1", "double_well", "morse3", "morse4"])
def test_mapping_trajectories_match_oracle(oracle, key, name):
    from nqcd_mi355x import _lib as L
    kind, params, om, n, mass, xm, xs, vm, vs, dt = _case(L, oracle, name)
    ntraj, nsteps, se = 256, 400, 40
    rng = np.random.default_rng(7)
    r0 = xm + xs * rng.standard_normal(ntraj)
    v0 = vm + vs * rng.standard_normal(ntraj)
    state = (np.arange(ntraj) % n).astype(np.int32)
    print(f"spin mapping parity: model={name} n={n} ntraj={ntraj} steps={nsteps}", flush=True)
    dev = _device(L, kind, params, n, mass, dt, r0, v0, state, nsteps, se, key)
    X0, P0 = oracle.mapping_sample(n, state, key)
    c = oracle.mapping_run(om, n, mass, r0, v0, X0, P0, dt=dt, nsteps=nsteps, save_every=se)
    assert np.allclose(dev["SAVE_POP_DI"][:, 0], (X0 ** 2 + P0 ** 2 - oracle.mapping_gamma(n)) / 2, rtol=0, atol=1e-14)
    assert np.allclose(dev["R"][:, 0], c["x"], rtol=1e-9, atol=1e-9)
    assert np.allclose(dev["V"][:, 0], c["v"], rtol=1e-9, atol=1e-11)
    assert np.allclose(dev["MAP_X"], c["X"], rtol=0, atol=1e-9)
    assert np.allclose(dev["MAP_P"], c["P"], rtol=0, atol=1e-9)
    assert np.allclose(dev["SAVE_POP_DI"], c["pop"], rtol=0, atol=1e-9)
    assert np.allclose(dev["SAVE_EKIN"], c["ekin"], rtol=1e-9, atol=1e-13)
    assert np.allclose(dev["SAVE_EPOT"], c["epot"], rtol=1e-9, atol=1e-12)
    # the populations moved (the run is not trivially adiabatic) and the device reductions agree with the fields
    assert np.abs(c["pop"][:, -1] - c["pop"][:, 0]).max() > 1e-3
    assert np.allclose(dev["red_pop"], dev["SAVE_POP_DI"].mean(axis=0), rtol=1e-12, atol=1e-15)
    assert np.array_equal(dev["noalloc_pop"], dev["red_pop"])
    assert np.allclose(dev["red_ekin"], dev["SAVE_EKIN"].sum(axis=0), rtol=1e-12, atol=1e-15)
    assert np.allclose(dev["red_etot"], (dev["SAVE_EKIN"] + dev["SAVE_EPOT"]).mean(axis=0), rtol=1e-12, atol=1e-15)
    # PopulationCorrelationFunction(sim, Diabatic()) (TimeCorrelationFunctions.jl:26-40,61-88): K(0) P(t)^T
    want = np.einsum("ta,tsb->sba", c["init"], c["pop"]) / ntraj
    assert np.allclose(dev["red_corr"], want, rtol=0, atol=1e-9)
    assert np.array_equal(dev["noalloc_corr"], dev["red_corr"])


def test_mapping_uploaded_variables(oracle, key):
    """nqcd_mi355x_load_mapping: uploaded mapping variables (not on the sampling circle) are integrated as
    given; X = P = None returns to the drawn ones."""
    from nqcd_mi355x import _lib as L
    kind, params, om, n, mass, xm, xs, vm, vs, dt = _case(L, oracle, "morse3")
    ntraj, nsteps = 128, 200
    rng = np.random.default_rng(11)
    r0 = xm + xs * rng.standard_normal(ntraj)
    v0 = vm + vs * rng.standard_normal(ntraj)
    state = np.zeros(ntraj, np.int32)
    X0, P0 = rng.standard_normal((ntraj, n)), rng.standard_normal((ntraj, n))
    print("spin mapping: uploaded mapping variables", flush=True)
    dev = _device(L, kind, params, n, mass, dt, r0, v0, state, nsteps, 0, key, X=X0, P=P0)
    c = oracle.mapping_run(om, n, mass, r0, v0, X0, P0, dt=dt, nsteps=nsteps)
    assert np.allclose(dev["R"][:, 0], c["x"], rtol=1e-9, atol=1e-9)
    assert np.allclose(dev["MAP_X"], c["X"], rtol=0, atol=1e-9)
    assert np.allclose(dev["MAP_P"], c["P"], rtol=0, atol=1e-9)
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_SPIN_MAPPING, n, 1, 1, [mass], dt, algorithm=L.ALG_MINT)
    E.set_model(0, kind, params)
    E.load_initial(r0.reshape(-1, 1), v0.reshape(-1, 1), None, state=state)
    E.load_mapping(X0, P0)
    E.load_mapping(None, None)
    E.run(0.0, 0.0, key)
    Xd = E.download(L.FIELD_MAP_X)
    E.close()
    Xs, _ = oracle.mapping_sample(n, state, key)
    assert np.allclose(Xd, Xs, rtol=0, atol=1e-15)


def test_mapping_large_ensemble_invariants(oracle, key):
    """1e5 Tully-1 trajectories through the avoided crossing: each trajectory's total population
    sum_i (X_i^2 + P_i^2 - gamma) / 2 stays 1 (the MInt rotation is orthogonal), its energy is conserved to
    2e-5 of the kinetic energy at dt = 1, and 256 trajectories picked across the ensemble match the oracle."""
    from nqcd_mi355x import _lib as L
    kind, params, om, n, mass, xm, xs, vm, vs, dt = _case(L, oracle, "tully1")
    ntraj, nsteps, se = 100_000, 1000, 100
    rng = np.random.default_rng(3)
    r0 = -6.0 + 0.5 * rng.standard_normal(ntraj)
    v0 = (15.0 + rng.standard_normal(ntraj)) / mass
    state = np.zeros(ntraj, np.int32)
    print(f"spin mapping invariants: ntraj={ntraj} steps={nsteps}", flush=True)
    dev = _device(L, kind, params, n, mass, dt, r0, v0, state, nsteps, se, key)
    pop = dev["SAVE_POP_DI"]
    assert np.allclose(pop.sum(axis=2), 1.0, rtol=0, atol=1e-10)
    e = dev["SAVE_EKIN"] + dev["SAVE_EPOT"]
    assert np.all(np.abs(e - e[:, :1]) <= 2e-5 * dev["SAVE_EKIN"][:, :1])
    assert dev["red_pop"][-1, 1] > 0.05   # population transferred through the crossing
    pick = np.linspace(0, ntraj - 1, 256).astype(np.int64)
    X0, P0 = oracle.mapping_sample(n, state[pick], key, traj_ids=pick)
    c = oracle.mapping_run(om, n, mass, r0[pick], v0[pick], X0, P0, dt=dt, nsteps=nsteps, save_every=se)
    assert np.allclose(pop[pick], c["pop"], rtol=0, atol=1e-9)
    assert np.allclose(dev["R"][pick, 0], c["x"], rtol=1e-9, atol=1e-9)


def test_mapping_run_dynamics(oracle, key):
    """Simulation[SpinMappingW] through run_dynamics: OutputDiabaticPopulation and OutputTotalEnergy under
    MeanReduction (device reductions) and OutputFinal per trajectory, against the oracle from the same
    initial conditions and draws."""
    import nqcd_mi355x as N
    from nqcd_mi355x import _lib as L
    ntraj, dt, tf = 512, 1.0, 600.0
    sim = N.Simulation[N.SpinMappingW](N.Atoms(2000.0), N.TullyModelOne())
    assert np.isclose(sim.method.gamma, oracle.mapping_gamma(2))
    rng = np.random.default_rng(5)
    r0 = -5.0 + 0.3 * rng.standard_normal(ntraj)
    v0 = np.full(ntraj, 20.0 / 2000.0)
    dist = N.DynamicalDistribution(list(v0.reshape(-1, 1, 1)), list(r0.reshape(-1, 1, 1)), (1, 1)) * \
        N.PureState(2, N.Diabatic())
    alg = N.EnsembleMI355X()
    print("spin mapping run_dynamics", flush=True)
    mean = N.run_dynamics(sim, (0.0, tf), dist, output=(N.OutputDiabaticPopulation, N.OutputTotalEnergy),
                          trajectories=ntraj, dt=dt, saveat=100.0, reduction=N.MeanReduction(), ensemble_algorithm=alg)
    per = N.run_dynamics(sim, (0.0, tf), dist, output=N.OutputFinal, trajectories=ntraj, dt=dt,
                         reduction=N.SortByTrajectoryReduction(), ensemble_algorithm=alg)
    state = np.ones(ntraj, np.int32)
    X0, P0 = oracle.mapping_sample(2, state, alg.seed)
    om = oracle.make_model(oracle.MODEL_TULLY1, 2, 1, TULLY1)
    c = oracle.mapping_run(om, 2, 2000.0, r0, v0, X0, P0, dt=dt, nsteps=int(tf / dt), save_every=100)
    assert np.allclose(np.array(mean["OutputDiabaticPopulation"]), c["pop"].mean(axis=0), rtol=0, atol=1e-10)
    assert np.allclose(mean["OutputTotalEnergy"], (c["ekin"] + c["epot"]).mean(axis=0), rtol=1e-10, atol=1e-14)
    assert np.allclose(np.array([p["OutputFinal"]["qmap"] for p in per]), c["X"], rtol=0, atol=1e-9)
    assert np.allclose(np.array([p["OutputFinal"]["r"][0, 0] for p in per]), c["x"], rtol=1e-9, atol=1e-9)


def test_mapping_run_dynamics_sampled(oracle, key):
    """run_dynamics with a Gaussian nuclear distribution: positions and velocities sampled on the device (the
    Philox Gaussian stream the oracle reproduces), mapping variables drawn in the run; the mean diabatic
    populations per save against the oracle from the same draws, and the per-save total population 1."""
    import nqcd_mi355x as N
    ntraj, dt, tf = 2048, 1.0, 800.0
    sim = N.Simulation[N.SpinMappingW](N.Atoms(2000.0), N.TullyModelOne())
    dist = N.DynamicalDistribution(N.Normal(12.0 / 2000.0, 1.0 / 2000.0), N.Normal(-3.0, 0.5), (1, 1)) * \
        N.PureState(1, N.Diabatic())
    alg = N.EnsembleMI355X()
    print("spin mapping run_dynamics, sampled", flush=True)
    res = N.run_dynamics(sim, (0.0, tf), dist, output=N.OutputDiabaticPopulation, trajectories=ntraj, dt=dt,
                         saveat=50.0, reduction=N.MeanReduction(), ensemble_algorithm=alg)
    pop = np.array(res["OutputDiabaticPopulation"])
    assert np.allclose(pop[0], [1.0, 0.0], atol=1e-14)
    assert np.allclose(pop.sum(axis=1), 1.0, atol=1e-12)
    r0, v0 = oracle.sample_gaussian(ntraj, 1, -3.0, 0.5, 12.0 / 2000.0, 1.0 / 2000.0, alg.seed)
    X0, P0 = oracle.mapping_sample(2, np.zeros(ntraj, np.int32), alg.seed)
    om = oracle.make_model(oracle.MODEL_TULLY1, 2, 1, TULLY1)
    c = oracle.mapping_run(om, 2, 2000.0, r0[:, 0], v0[:, 0], X0, P0, dt=dt, nsteps=int(tf / dt), save_every=50)
    assert np.allclose(pop, c["pop"].mean(axis=0), rtol=0, atol=1e-10)
    assert abs(pop[-1, 1]) > 0.05


@pytest.mark.parametrize("name,gamma", [("tully1", 0.5), ("double_well", 0.3), ("morse3", 0.1), ("morse4", 0.0)])
def test_ecmm_trajectories_match_oracle(oracle, key, name, gamma):
    """eCMM (cmm.jl) by MInt: the device's sphere draws equal the oracle's, and trajectories, populations (the
    inverse mapping kernel) and energies agree within 1e-9 over 400 steps; on the double well (tr dV != 0) this
    includes MInt's trace term at gamma / 2 of eCMM's (mint.jl:85), reproduced on both sides."""
    from nqcd_mi355x import _lib as L
    kind, params, om, n, mass, xm, xs, vm, vs, dt = _case(L, oracle, name)
    ntraj, nsteps, se = 256, 400, 40
    rng = np.random.default_rng(9)
    r0 = xm + xs * rng.standard_normal(ntraj)
    v0 = vm + vs * rng.standard_normal(ntraj)
    print(f"eCMM parity: model={name} n={n} gamma={gamma}", flush=True)
    dev = _device(L, kind, params, n, mass, dt, r0, v0, None, nsteps, se, key, ecmm_gamma=gamma)
    X0, P0 = oracle.ecmm_sample(n, ntraj, gamma, key)
    c = oracle.mapping_run(om, n, mass, r0, v0, X0, P0, dt=dt, nsteps=nsteps, save_every=se, method=oracle.MAP_ECMM,
                           gamma=gamma)
    R2 = 2 + 2 * n * gamma
    assert np.isfinite(c["pop"]).all()
    assert np.allclose(dev["R"][:, 0], c["x"], rtol=1e-9, atol=1e-9)
    assert np.allclose(dev["V"][:, 0], c["v"], rtol=1e-9, atol=1e-11)
    assert np.allclose(dev["MAP_X"], c["X"], rtol=0, atol=1e-9)
    assert np.allclose(dev["MAP_P"], c["P"], rtol=0, atol=1e-9)
    assert np.allclose((dev["MAP_X"] ** 2 + dev["MAP_P"] ** 2).sum(1), R2, rtol=1e-12)
    assert np.allclose(dev["SAVE_POP_DI"], c["pop"], rtol=0, atol=1e-9)
    assert np.allclose(dev["SAVE_EKIN"], c["ekin"], rtol=1e-9, atol=1e-13)
    assert np.allclose(dev["SAVE_EPOT"], c["epot"], rtol=1e-9, atol=1e-12)
    assert np.allclose(dev["red_pop"], dev["SAVE_POP_DI"].mean(axis=0), rtol=1e-12, atol=1e-15)
    # eCMM's correlation: normalisation F, initial estimator the mapping kernel (cmm.jl:106-119)
    want = n * np.einsum("ta,tsb->sba", c["init"], c["pop"]) / ntraj
    assert np.allclose(dev["red_corr"], want, rtol=0, atol=1e-9)


def test_ecmm_run_dynamics(oracle, key):
    """Simulation[eCMM](atoms, model, gamma=0.5) through run_dynamics with a Gaussian distribution and any
    electronic state: the mean diabatic populations against the oracle from the same draws."""
    import nqcd_mi355x as N
    ntraj, dt, tf, gamma = 1024, 1.0, 600.0, 0.5
    sim = N.Simulation[N.eCMM](N.Atoms(2000.0), N.TullyModelOne(), gamma=gamma)
    dist = N.DynamicalDistribution(N.Normal(12.0 / 2000.0, 1.0 / 2000.0), N.Normal(-3.0, 0.5), (1, 1)) * \
        N.PureState(1, N.Diabatic())
    alg = N.EnsembleMI355X()
    print("eCMM run_dynamics", flush=True)
    res = N.run_dynamics(sim, (0.0, tf), dist, output=(N.OutputDiabaticPopulation, N.OutputTotalEnergy),
                         trajectories=ntraj, dt=dt, saveat=100.0, reduction=N.MeanReduction(), ensemble_algorithm=alg)
    r0, v0 = oracle.sample_gaussian(ntraj, 1, -3.0, 0.5, 12.0 / 2000.0, 1.0 / 2000.0, alg.seed)
    X0, P0 = oracle.ecmm_sample(2, ntraj, gamma, alg.seed)
    om = oracle.make_model(oracle.MODEL_TULLY1, 2, 1, TULLY1)
    c = oracle.mapping_run(om, 2, 2000.0, r0[:, 0], v0[:, 0], X0, P0, dt=dt, nsteps=int(tf / dt), save_every=100,
                           method=oracle.MAP_ECMM, gamma=gamma)
    assert np.allclose(np.array(res["OutputDiabaticPopulation"]), c["pop"].mean(axis=0), rtol=0, atol=1e-10)
    assert np.allclose(res["OutputTotalEnergy"], (c["ekin"] + c["epot"]).mean(axis=0), rtol=1e-10, atol=1e-14)


def test_mapping_population_correlation_paths(oracle, key):
    """PopulationCorrelationFunction(sim, Diabatic()) through run_dynamics for eCMM: the device MeanReduction equals
    the mean of the per-trajectory values (SortByTrajectoryReduction, built from the downloaded initial estimator
    and saves) and, at t0, F times the mean of K(0) P(0)^T."""
    import nqcd_mi355x as N
    ntraj, dt, tf = 512, 1.0, 300.0
    sim = N.Simulation[N.eCMM](N.Atoms(2000.0), N.TullyModelOne(), gamma=0.3)
    dist = N.DynamicalDistribution(N.Normal(12.0 / 2000.0, 1.0 / 2000.0), N.Normal(-2.0, 0.5), (1, 1)) * \
        N.PureState(1, N.Diabatic())
    corr = N.PopulationCorrelationFunction(sim, N.Diabatic())
    alg = N.EnsembleMI355X()
    print("eCMM population correlation", flush=True)
    mean = N.run_dynamics(sim, (0.0, tf), dist, output=corr, trajectories=ntraj, dt=dt, saveat=50.0,
                          reduction=N.MeanReduction(), ensemble_algorithm=alg)
    per = N.run_dynamics(sim, (0.0, tf), dist, output=corr, trajectories=ntraj, dt=dt, saveat=50.0,
                         reduction=N.SortByTrajectoryReduction(), ensemble_algorithm=alg)
    host = np.mean([np.array(p["PopulationCorrelationFunction"]) for p in per], axis=0)
    dev = np.array(mean["PopulationCorrelationFunction"])
    assert dev.shape == host.shape == (7, 2, 2)
    assert np.allclose(dev, host, rtol=1e-12, atol=1e-14)
    r0, v0 = oracle.sample_gaussian(ntraj, 1, -2.0, 0.5, 12.0 / 2000.0, 1.0 / 2000.0, alg.seed)
    X0, P0 = oracle.ecmm_sample(2, ntraj, 0.3, alg.seed)
    om = oracle.make_model(oracle.MODEL_TULLY1, 2, 1, TULLY1)
    c = oracle.mapping_run(om, 2, 2000.0, r0[:, 0], v0[:, 0], X0, P0, dt=dt, nsteps=int(tf / dt), save_every=50,
                           method=oracle.MAP_ECMM, gamma=0.3)
    want = 2 * np.einsum("ta,tsb->sab", c["init"], c["pop"]) / ntraj
    assert np.allclose(dev, want, rtol=0, atol=1e-9)
