"""GPU parity of the two-state FSSH / Ehrenfest kernels against the FP64 oracle.

Bar (north star): identical hop sequences, per-trajectory R/P/sigma within 1e-9
relative on the FP64 path over short windows, ensemble populations within 2 sigma;
the FP32-nuclei path within 0.01 absolute final population of the FP64 oracle.
All calls go through libnqcd_mi355x.so's C ABI.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TULLY = {0: [0.01, 1.6, 0.005, 1.0], 1: [0.1, 0.28, 0.015, 0.06, 0.05], 2: [6e-4, 0.1, 0.9]}
MASS = 2000.0


def _ens():
    from nqcd_mi355x import _lib as L
    return L, L.Ensemble(0)


def _window_inputs(oracle, key, ntraj, kind):
    """Initial conditions near the coupling region so hops happen inside the window."""
    rng = np.random.default_rng(11 + kind)
    x0 = -6.5 if kind == 2 else -2.0  # Tully 3 couples strongly near x ~ -5.7
    r0 = rng.normal(x0, 0.6, size=(ntraj, 1))
    p0 = rng.uniform(6.0, 32.0, size=ntraj)
    v0 = (p0 / MASS)[:, None]
    sig0 = np.zeros((ntraj, 2, 2), complex)
    start = rng.integers(0, 2, size=ntraj)
    sig0[np.arange(ntraj), start, start] = 1.0
    # a quarter start in a coherent superposition (exercises Ehrenfest and the off-diagonals)
    q = ntraj // 4
    c = np.stack([np.full(q, np.sqrt(0.7)), np.sqrt(0.3) * np.exp(1j * rng.uniform(0, 2 * np.pi, q))], 1)
    sig0[:q] = np.einsum("ti,tj->tij", c, c.conj())
    return r0, v0, sig0, start.astype(np.int32)


def _run_gpu(L, E, method, kind, r0, v0, sig0, state0, nsteps, key, precision=0, schedule=0, save_every=0, dt=1.0,
             term=None, rescaling=0):
    E.set_problem(method, 2, 1, 1, [MASS], dt, precision=precision, schedule=schedule, rescaling=rescaling)
    E.set_model(0, kind, TULLY[kind])
    E.load_initial(r0, v0, sig0, state=state0)
    st = E.run(0.0, nsteps * dt, key, save_every=save_every, term=term)
    out = {"r": E.download(L.FIELD_R), "v": E.download(L.FIELD_V),
           "sig": E.download(L.FIELD_SIGMA_RE) + 1j * E.download(L.FIELD_SIGMA_IM),
           "state": E.download(L.FIELD_STATE), "steps": E.download(L.FIELD_STEPS), "nhops": E.download(L.FIELD_NHOPS)}
    if save_every:
        out["save_state"] = E.download(L.FIELD_SAVE_STATE)
        out["pop_ad"] = E.download(L.FIELD_SAVE_POP_AD)
        out["pop_di"] = E.download(L.FIELD_SAVE_POP_DI)
    out["stats"] = st
    return out


def _run_cpu(oracle, method, kind, r0, v0, sig0, state0, nsteps, key, save_every=0, dt=1.0, term_x=0.0, rescaling=0):
    m = oracle.make_model(kind, 2, 1, TULLY[kind])
    nsave = nsteps // save_every + 1 if save_every else 1
    return oracle.run_ensemble(method, m, MASS, r0, v0, sig0, state0=state0, dt=dt, nsteps=nsteps, key=key,
                               save_every=save_every, nsave=nsave, term_x=term_x, rescaling=rescaling)


def _assert_window_parity(g, c, rtol=1e-9):
    assert np.array_equal(g["state"], c["state"])
    assert np.array_equal(g["nhops"], c["nhops"])
    assert np.allclose(g["r"], c["r"], rtol=rtol, atol=1e-12)
    assert np.allclose(g["v"], c["v"], rtol=rtol, atol=1e-14)
    sc = c["sig_re"] + 1j * c["sig_im"]
    assert np.allclose(g["sig"], sc, rtol=rtol, atol=1e-11)


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("schedule", [0, 1])
def test_fssh_fp64_window_parity(oracle, key, kind, schedule):
    """1024 trajectories, 400 steps: identical hop sequences, R/P/sigma to 1e-9 relative."""
    L, E = _ens()
    ntraj, nsteps = 1024, 400
    r0, v0, sig0, st0 = _window_inputs(oracle, key, ntraj, kind)
    c = _run_cpu(oracle, 0, kind, r0, v0, sig0, st0, nsteps, key, save_every=1)
    g = _run_gpu(L, E, 0, kind, r0, v0, sig0, st0, nsteps, key, schedule=schedule, save_every=1)
    assert c["nhops"].sum() > 20, "window must contain hops"
    assert np.array_equal(g["save_state"], c["save_state"])  # hop sequence, step by step
    _assert_window_parity(g, c)
    assert np.allclose(g["pop_di"], c["pop_di"], rtol=1e-9, atol=1e-11)
    assert np.all(g["steps"] == nsteps)
    E.close()


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_ehrenfest_fp64_window_parity(oracle, key, kind):
    L, E = _ens()
    ntraj, nsteps = 512, 400
    r0, v0, sig0, _ = _window_inputs(oracle, key, ntraj, kind)
    c = _run_cpu(oracle, 1, kind, r0, v0, sig0, None, nsteps, key, save_every=10)
    g = _run_gpu(L, E, 1, kind, r0, v0, sig0, None, nsteps, key, save_every=10)
    assert np.allclose(g["r"], c["r"], rtol=1e-9, atol=1e-12)
    assert np.allclose(g["v"], c["v"], rtol=1e-9, atol=1e-14)
    assert np.allclose(g["sig"], c["sig_re"] + 1j * c["sig_im"], rtol=1e-9, atol=1e-11)
    assert np.allclose(g["pop_ad"], c["pop_ad"], rtol=1e-9, atol=1e-11)
    assert np.allclose(g["pop_di"], c["pop_di"], rtol=1e-9, atol=1e-11)
    E.close()


def test_rescaling_modes_parity(oracle, key):
    L, E = _ens()
    ntraj, nsteps = 1024, 400
    r0, v0, sig0, st0 = _window_inputs(oracle, key, ntraj, 1)
    v0 = v0 * 0.35  # slow: many frustrated hops
    for resc in (1, 2):
        c = _run_cpu(oracle, 0, 1, r0, v0, sig0, st0, nsteps, key, rescaling=resc)
        g = _run_gpu(L, E, 0, 1, r0, v0, sig0, st0, nsteps, key, rescaling=resc)
        _assert_window_parity(g, c)
    E.close()


def test_tully1_workload_fp64_matches_oracle(oracle, key):
    """The 1000-trajectory Tully-1 workload (p0 = 20, 3000 steps of dt = 1): same final states as
    the oracle, hence the same populations and transmission/reflection probabilities."""
    L, E = _ens()
    ntraj, p0 = 1000, 20.0
    r0, v0 = oracle.sample_gaussian(ntraj, 1, -5.0, p0 / (20 * np.sqrt(2)), p0 / MASS, 0.0, key)
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    st0 = np.zeros(ntraj, np.int32)
    c = _run_cpu(oracle, 0, 0, r0, v0, sig0, st0, 3000, key)
    g = _run_gpu(L, E, 0, 0, r0, v0, sig0, st0, 3000, key)
    mismatch = np.mean(g["state"] != c["state"])
    assert mismatch <= 0.002
    pc = np.bincount(c["state"], minlength=2) / ntraj
    sig = np.sqrt(pc * (1 - pc) / ntraj)
    assert np.all(np.abs(np.bincount(g["state"], minlength=2) / ntraj - pc) <= 2 * sig + 1e-12)
    scat = E.reduce(L.OUT_SCATTERING_1D, L.RED_MEAN, 4)
    x = c["r"][:, 0]
    ref = [np.mean((x <= 0) & (c["state"] == 0)), np.mean((x <= 0) & (c["state"] == 1)),
           np.mean((x > 0) & (c["state"] == 0)), np.mean((x > 0) & (c["state"] == 1))]
    assert np.allclose(scat, ref, atol=2.5 * 0.5 / np.sqrt(ntraj))
    fin = E.reduce(L.OUT_FINAL_ADIABATIC_POPULATION, L.RED_MEAN, 2)
    assert np.allclose(fin, np.bincount(g["state"], minlength=2) / ntraj, atol=1e-14)
    E.close()


def test_fp32_nuclei_population_error(oracle, key):
    """FP32 nuclei / FP64 electronics: final populations within 0.01 of the FP64 oracle."""
    L, E = _ens()
    ntraj = 4000
    res = {}
    for kind in (0, 1, 2):
        p0 = 16.0
        r0, v0 = oracle.sample_gaussian(ntraj, 1, -5.0, p0 / (20 * np.sqrt(2)), p0 / MASS, 0.0, key)
        sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
        st0 = np.zeros(ntraj, np.int32)
        c = _run_cpu(oracle, 0, kind, r0, v0, sig0, st0, 10000, key, term_x=10.0)
        g = _run_gpu(L, E, 0, kind, r0, v0, sig0, st0, 10000, key, precision=1, term=(0, 10.0))
        pc = np.bincount(c["state"], minlength=2) / ntraj
        pg = np.bincount(g["state"], minlength=2) / ntraj
        res[kind] = (pc, pg)
        assert np.abs(pg - pc).max() <= 0.01, (kind, pc, pg)
    E.close()


def test_device_sampler_matches_oracle(oracle, key):
    """TULLY_GRID device sampling = oracle Philox normals (x0 = -5, p0 grid 5..35, three models)."""
    L, E = _ens()
    ntraj = 3 * 31 * 4
    E.set_problem(0, 2, 1, 1, [MASS], 1.0)
    for k in range(3):
        E.set_model(k, k, TULLY[k])
    E.sample_initial(L.SAMPLER_TULLY_GRID, ntraj, [-5.0, 5.0, 31], key, n_models=3)
    E.run(0.0, 0.0, key)  # zero steps: outputs equal the initial conditions
    ids = np.arange(ntraj)
    p0 = 5.0 + (ids // 3) % 31
    r0, v0 = oracle.sample_gaussian(ntraj, 1, -5.0, (p0 / (20 * np.sqrt(2)))[:, None], (p0 / MASS)[:, None], 0.0, key)
    assert np.allclose(E.download(L.FIELD_R), r0, rtol=1e-13, atol=1e-14)
    assert np.allclose(E.download(L.FIELD_V), v0, rtol=1e-15)
    E.close()


def test_mixed_models_uneven_lengths_parity(oracle, key):
    """Mixed Tully 1/2/3 ensemble with |x|>10 termination: persistent refill queue gives the
    oracle's per-trajectory results (FP64)."""
    L, E = _ens()
    ntraj = 3 * 31 * 8
    E.set_problem(0, 2, 1, 1, [MASS], 1.0)
    for k in range(3):
        E.set_model(k, k, TULLY[k])
    E.sample_initial(L.SAMPLER_TULLY_GRID, ntraj, [-5.0, 5.0, 31], key, n_models=3)
    E.run(0.0, 10000.0, key, term=(0, 10.0))
    g = {f: E.download(getattr(L, "FIELD_" + f.upper())) for f in ("r", "v", "state", "steps", "nhops")}
    ids = np.arange(ntraj)
    p0 = 5.0 + (ids // 3) % 31
    r0, v0 = oracle.sample_gaussian(ntraj, 1, -5.0, (p0 / (20 * np.sqrt(2)))[:, None], (p0 / MASS)[:, None], 0.0, key)
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    models = [oracle.make_model(k, 2, 1, TULLY[k]) for k in range(3)]
    c = oracle.run_ensemble(0, models, MASS, r0, v0, sig0, state0=np.zeros(ntraj, np.int32), model_id=ids % 3,
                            dt=1.0, nsteps=10000, key=key, term_x=10.0)
    same = (g["state"] == c["state"]) & (g["steps"] == c["steps"])
    assert same.mean() >= 0.995
    assert np.allclose(g["r"][same], c["r"][same], rtol=1e-8, atol=1e-9)
    st = E.stats()
    assert st["traj_steps"] == g["steps"].sum()
    assert len(np.unique(g["steps"])) > 100  # uneven lengths
    E.close()


def test_abi_errors_are_reported():
    L, E = _ens()
    with pytest.raises(L.NQCDError) as ei:
        E.run(0.0, 1.0, 0)
    assert ei.value.code == L.ERR_STATE
    with pytest.raises(L.NQCDError) as ei:
        E.set_problem(0, 2, 1, 1, [MASS], 1.0, rescaling=7)
    assert "rescaling" in str(ei.value)
    with pytest.raises(L.NQCDError) as ei:  # IESH integrates with VerletwithElectronics only
        E.set_problem(L.METHOD_IESH, 60, 1, 1, [MASS], 1.0, algorithm=L.ALG_BAB_TSIT5, n_electrons=30)
    assert ei.value.code == L.ERR_UNSUPPORTED
    with pytest.raises(L.NQCDError) as ei:
        E.set_problem(L.METHOD_IESH, 60, 1, 1, [MASS], 1.0, algorithm=L.ALG_VERLET_ELECTRONICS)
    assert ei.value.code == L.ERR_ARG and "n_electrons" in str(ei.value)
    E.close()


@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("predicate,value", [(2, 2.0), (3, -7.0), (1, 6.0)])
def test_termination_predicates_parity(oracle, key, schedule, predicate, value):
    """TerminatingCallback predicates evaluated on the device (x > 2, x < -7, |x| > 6) in the persistent
    and the step-per-launch kernels against the oracle: identical step counts, states and hops, final
    positions within 1e-9 (trajectories moving both ways, so both one-sided predicates fire)."""
    L, E = _ens()
    rng = np.random.default_rng(40 + predicate)
    ntraj = 256
    r0 = rng.normal(-5.0, 0.5, (ntraj, 1))
    v0 = rng.choice([-1.0, 1.0], (ntraj, 1)) * rng.uniform(10.0, 30.0, (ntraj, 1)) / MASS
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    st0 = np.zeros(ntraj, np.int32)
    g = _run_gpu(L, E, 0, 0, r0, v0, sig0, st0, 3000, key, schedule=schedule, term=(0, value, predicate))
    m = oracle.make_model(0, 2, 1, TULLY[0])
    c = oracle.run_ensemble(0, m, MASS, r0, v0, sig0, state0=st0, dt=1.0, nsteps=3000, key=key, term_x=value,
                            term_kind=0 if predicate == 1 else predicate)
    assert np.array_equal(g["steps"], c["steps"])
    assert (g["steps"] < 3000).sum() > 20  # the predicate ended a good part of the ensemble
    _assert_window_parity(g, c)
    E.close()


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("schedule", [0, 1])
def test_energy_outputs_parity(oracle, key, method, schedule):
    """OutputKineticEnergy / OutputTotalEnergy (DynamicsOutputs.jl:90-98) per save slot from the device
    (persistent and step-per-launch kernels) against the oracle: per trajectory within 1e-9, and the
    device Mean reduction of the total against the oracle's mean."""
    L, E = _ens()
    ntraj = 128
    r0, v0, sig0, st0 = _window_inputs(oracle, key, ntraj, 0)
    if method == 0:
        sig0 = np.zeros((ntraj, 2, 2), complex); sig0[np.arange(ntraj), st0, st0] = 1
    g = _run_gpu(L, E, method, 0, r0, v0, sig0, st0, 400, key, schedule=schedule, save_every=40)
    ekin, epot = E.download(L.FIELD_SAVE_EKIN), E.download(L.FIELD_SAVE_EPOT)
    tot_mean = E.reduce(L.OUT_TOTAL_ENERGY, L.RED_MEAN, 11)
    kin_sum = E.reduce(L.OUT_KINETIC_ENERGY, L.RED_SUM, 11)
    E.close()
    m = oracle.make_model(0, 2, 1, TULLY[0])
    c = oracle.run_ensemble(method, m, MASS, r0, v0, sig0, state0=st0, dt=1.0, nsteps=400, key=key, save_every=40,
                            nsave=11, energies=True)
    assert np.array_equal(g["state"], c["state"])
    assert np.allclose(ekin, c["ekin"], rtol=1e-9, atol=1e-14)
    assert np.allclose(epot, c["epot"], rtol=1e-9, atol=1e-14)
    assert np.allclose(tot_mean, (c["ekin"] + c["epot"]).mean(axis=0), rtol=1e-9, atol=1e-14)
    assert np.allclose(kin_sum, c["ekin"].sum(axis=0), rtol=1e-9)


def test_autotune_launch_shapes_leave_results_unchanged(oracle, key, tmp_path, monkeypatch):
    """A shape read from the tuning cache file (64 threads, 3 workgroups per CU, for devices of any
    size) is the one launched; nqcd_mi355x_autotune then times the candidate shapes on a sample run
    and later runs take its choice.  Every shape gives bit-identical trajectories (ids, not lanes, key
    the random numbers)."""
    L, E = _ens()
    ntraj = 3000
    r0, v0, sig0, st0 = _window_inputs(oracle, key, ntraj, 0)
    ref = _run_gpu(L, E, 0, 0, r0, v0, sig0, st0, 300, key, precision=1)
    forced_file, default_file = tmp_path / "tune.txt", tmp_path / "tune_default.txt"
    forced_file.write_text("1 2 1 0 64 3\n")
    default_file.write_text("1 2 1 0 256 0\n")
    monkeypatch.setenv("NQCD_MI355X_TUNE_CACHE", str(forced_file))
    try:
        forced = _run_gpu(L, E, 0, 0, r0, v0, sig0, st0, 300, key, precision=1)
    finally:
        monkeypatch.setenv("NQCD_MI355X_TUNE_CACHE", str(default_file))  # leave the default shape in the table
        restored = _run_gpu(L, E, 0, 0, r0, v0, sig0, st0, 10, key, precision=1)
    assert forced["stats"]["block"] == 64 and 1 <= forced["stats"]["grid"] <= (ntraj + 63) // 64
    assert restored["stats"]["block"] == 256
    res = E.autotune(0.0, 100.0, key)
    assert res["family"] == 1 and res["n_candidates"] >= 6 and res["block"] in (64, 128, 256)
    assert 0 < res["best_ms"] <= res["default_ms"]
    tuned = _run_gpu(L, E, 0, 0, r0, v0, sig0, st0, 300, key, precision=1)
    E.close()
    assert tuned["stats"]["block"] == res["block"]
    for k in ("r", "v", "sig", "state", "steps", "nhops"):
        assert np.array_equal(tuned[k], ref[k]), k
        assert np.array_equal(forced[k], ref[k]), k
    assert ref["nhops"].sum() > 0


@pytest.mark.parametrize("schedule", [0, 1])
def test_trajectories_starting_past_the_boundary_take_one_step(oracle, key, schedule):
    """A third of the ensemble starts beyond |x| > 10: TerminatingCallback is evaluated after each step,
    never at t0, so those take exactly one step; the whole ensemble matches the oracle (hop sequences,
    lengths, FP64 state) on both schedules."""
    L, E = _ens()
    ntraj = 130
    r0, v0, sig0, st0 = _window_inputs(oracle, key, ntraj, 0)
    past = np.arange(ntraj) % 3 == 0
    r0[past, 0] = np.where(np.arange(past.sum()) % 2 == 0, 10.5, -11.0)
    g = _run_gpu(L, E, 0, 0, r0, v0, sig0, st0, 600, key, schedule=schedule, save_every=20, term=(0, 10.0))
    E.close()
    c = _run_cpu(oracle, 0, 0, r0, v0, sig0, st0, 600, key, save_every=20, term_x=10.0)
    assert np.all(g["steps"][past] == 1)
    assert np.array_equal(g["steps"], c["steps"]) and np.array_equal(g["state"], c["state"])
    assert np.array_equal(g["save_state"], c["save_state"])
    assert np.allclose(g["r"], c["r"], rtol=1e-9, atol=1e-12)
    assert np.allclose(g["v"], c["v"], rtol=1e-9, atol=1e-14)


@pytest.mark.parametrize("precision", [0, 1])
def test_set_aside_policies_leave_results_unchanged(oracle, key, monkeypatch, precision):
    """Once the queue has drained, waves with few running lanes set them aside and the next launch takes them
    up packed (NQCD_MI355X_SUSPEND = "rounds,threshold").  The state is carried bit for bit and the random
    numbers are keyed by trajectory id and step, so a single launch, the default policy, and every drained
    wave setting its lanes aside at every step give identical trajectories, saves and lengths; and the FP64
    path matches the oracle under the default policy."""
    L, E = _ens()
    ntraj = 20000
    r0, v0, sig0, st0 = _window_inputs(oracle, key, ntraj, 0)
    runs = {}
    for pol in ("1,0", None, "3,63", "8,64"):
        if pol is None:
            monkeypatch.delenv("NQCD_MI355X_SUSPEND", raising=False)
        else:
            monkeypatch.setenv("NQCD_MI355X_SUSPEND", pol)
        print(f"set-aside policy {pol}", flush=True)
        runs[pol] = _run_gpu(L, E, 0, 0, r0, v0, sig0, st0, 1500, key, precision=precision, save_every=50,
                             term=(0, 10.0))
    monkeypatch.delenv("NQCD_MI355X_SUSPEND", raising=False)
    E.close()
    ref = runs["1,0"]
    assert ref["stats"]["launches"] == 2 and runs["8,64"]["stats"]["launches"] == 16
    assert len(np.unique(ref["steps"])) > 50 and ref["nhops"].sum() > 0   # uneven lengths: a real tail
    for pol, g in runs.items():
        for k in ("r", "v", "sig", "state", "steps", "nhops", "save_state", "pop_ad", "pop_di"):
            assert np.array_equal(g[k], ref[k]), (pol, k)
    if precision == 0:
        sel = np.arange(0, ntraj, 40)
        m = oracle.make_model(0, 2, 1, TULLY[0])
        c = oracle.run_ensemble(0, m, MASS, r0[sel], v0[sel], sig0[sel], state0=st0[sel], dt=1.0, nsteps=1500, key=key,
                                term_x=10.0, traj_ids=sel)
        assert np.array_equal(runs[None]["steps"][sel], c["steps"]) and np.array_equal(runs[None]["state"][sel], c["state"])
        assert np.allclose(runs[None]["r"][sel], c["r"], rtol=1e-9, atol=1e-12)
