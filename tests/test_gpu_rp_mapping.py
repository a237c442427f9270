"""Ring-polymer mapping dynamics by RingPolymerMInt on the device (NRPMD, nrpmd.jl; ring-polymer eCMM, rpecmm.jl;
ringpolymer_mint.jl) against the FP64 oracle: mapping variables drawn on the device equal to the oracle's draws,
bead positions and velocities, mapping variables, per-save populations, energies and centroids within 1e-9 over
300 steps at 1 to 16 beads (including bead counts that leave lanes of a 64-lane workgroup idle) and two to four
states; uploaded mapping variables; at 1e5 ring polymers the population sum and every ring polymer's energy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DW = [1 / 1200, 1 / 40, 0.01, 0.005, 1.0]
TULLY1 = [0.01, 1.6, 0.005, 1.0]
TULLY2 = [0.1, 0.28, 0.015, 0.06, 0.05]
MORSE3 = [0.02, 0.5, 4.0, 0.0, 0.02, 0.5, 4.0, 0.004, 0.02, 0.5, 4.0, 0.008,
          0.003, 0.5, 4.0, 0.002, 0.5, 4.2, 0.003, 0.5, 4.4]
MORSE4 = [0.02, 0.5, 4.0, 0.0, 0.02, 0.5, 4.0, 0.004, 0.02, 0.5, 4.0, 0.008, 0.02, 0.5, 4.0, 0.012,
          0.003, 0.5, 4.0, 0.002, 0.5, 4.2, 0.001, 0.5, 4.4, 0.003, 0.5, 4.1, 0.002, 0.5, 4.3, 0.003, 0.5, 4.5]

# name: (method, gamma, device kind, params, oracle kind, oracle params, n, P, mass, T, x0, spread, dt)
CASES = {
    "nrpmd_dw_P4": ("NRPMD", 0.5, "DW_GAUSSIAN", DW, "MODEL_DW_GAUSSIAN", DW, 2, 4, 1836.0, 1 / 8, -0.5, 0.3, 1.0),
    "nrpmd_tully2_P10": ("NRPMD", 0.5, "TULLY2", TULLY2, "MODEL_TULLY2", TULLY2, 2, 10, 2000.0, 1e-3, -1.0, 0.05, 2.0),
    "ecmm_morse3_P3": ("ECMM", 0.3, "MORSE_N", [3] + MORSE3, "MODEL_MORSE_N", MORSE3, 3, 3, 20000.0, 1e-3, 4.1, 0.05, 2.0),
    "ecmm_tully1_P16": ("ECMM", 0.0, "TULLY1", TULLY1, "MODEL_TULLY1", TULLY1, 2, 16, 2000.0, 1e-3, -2.0, 0.05, 1.0),
    "nrpmd_morse4_P1": ("NRPMD", 0.25, "MORSE_N", [4] + MORSE4, "MODEL_MORSE_N", MORSE4, 4, 1, 20000.0, 1e-3, 4.1, 0.1, 2.0),
    "ecmm_morse4_P8": ("ECMM", 0.1, "MORSE_N", [4] + MORSE4, "MODEL_MORSE_N", MORSE4, 4, 8, 20000.0, 1e-3, 4.1, 0.05, 2.0),
}


def _inputs(case, ntraj, seed):
    method, g, dk, dp, ok, op, n, P, mass, T, x0, spread, dt = case
    rng = np.random.default_rng(seed)
    r0 = x0 + spread * rng.standard_normal((ntraj, P))
    v0 = (2e-3 if mass < 5000 else 2e-4) + 1e-4 * rng.standard_normal((ntraj, P))
    return r0, v0


def _device(L, case, r0, v0, state, nsteps, se, key, X=None, P_=None):
    method, g, dk, dp, ok, op, n, P, mass, T, x0, spread, dt = case
    E = L.Ensemble(0)
    E.set_problem(getattr(L, "METHOD_" + method), n, 1, 1, [mass], dt, algorithm=L.ALG_RP_MINT, nbeads=P,
                  temperature=T, mapping_gamma=g)
    E.set_model(0, getattr(L, "MODEL_" + dk), dp)
    E.load_initial(r0, v0, None, state=state)
    if X is not None:
        E.load_mapping(X, P_)
    E.run(0.0, dt * nsteps, key, save_every=se)
    out = {f: E.download(getattr(L, "FIELD_" + f)) for f in ("R", "V", "MAP_X", "MAP_P", "SAVE_POP_DI", "SAVE_EKIN",
                                                             "SAVE_EPOT", "SAVE_CENTROID", "STEPS")}
    out["red_pop"] = E.reduce(L.OUT_DIABATIC_POPULATION, L.RED_MEAN, n * E.nsave).reshape(E.nsave, n)
    out["red_etot"] = E.reduce(L.OUT_TOTAL_ENERGY, L.RED_SUM, E.nsave)
    out["red_cen"] = E.reduce(L.OUT_CENTROID_POSITION, L.RED_MEAN, E.nsave)
    out["red_corr"] = E.reduce(L.OUT_POPULATION_CORRELATION, L.RED_MEAN, n * n * E.nsave).reshape(E.nsave, n, n)
    out["norm"] = n if method == "ECMM" else 1
    E.close()
    return out


def _oracle(O, case, r0, v0, q0, p0, nsteps, se):
    method, g, dk, dp, ok, op, n, P, mass, T, x0, spread, dt = case
    m = O.make_model(getattr(O, ok), n, 1, op)
    om = O.RPMAP_NRPMD if method == "NRPMD" else O.RPMAP_ECMM
    return O.rpmap_run(om, m, n, P, mass, T, g, r0, v0, q0, p0, dt=dt, nsteps=nsteps, save_every=se)


def _check(dev, c, nsteps):
    assert np.isfinite(c["pop"]).all()
    assert np.all(dev["STEPS"] == nsteps)
    assert np.allclose(dev["R"], c["r"], rtol=1e-9, atol=1e-9)
    assert np.allclose(dev["V"], c["v"], rtol=1e-9, atol=1e-11)
    assert np.allclose(dev["MAP_X"], c["q"], rtol=0, atol=1e-9)
    assert np.allclose(dev["MAP_P"], c["p"], rtol=0, atol=1e-9)
    assert np.allclose(dev["SAVE_POP_DI"], c["pop"], rtol=0, atol=1e-9)
    assert np.allclose(dev["SAVE_EKIN"], c["ekin"], rtol=1e-9, atol=1e-13)
    assert np.allclose(dev["SAVE_EPOT"], c["epot"], rtol=1e-9, atol=1e-11)
    assert np.allclose(dev["SAVE_CENTROID"], c["centroid"], rtol=1e-9, atol=1e-9)
    assert np.allclose(dev["red_pop"], dev["SAVE_POP_DI"].mean(axis=0), rtol=1e-12, atol=1e-15)
    assert np.allclose(dev["red_etot"], (dev["SAVE_EKIN"] + dev["SAVE_EPOT"]).sum(axis=0), rtol=1e-12, atol=1e-13)
    assert np.allclose(dev["red_cen"], dev["SAVE_CENTROID"].mean(axis=0), rtol=1e-12, atol=1e-14)
    want = dev["norm"] * np.einsum("ta,tsb->sba", c["init"], c["pop"]) / len(c["init"])
    assert np.allclose(dev["red_corr"], want, rtol=0, atol=1e-9)


@pytest.mark.parametrize("name", list(CASES))
def test_rp_mapping_matches_oracle(oracle, key, name):
    from nqcd_mi355x import _lib as L
    case = CASES[name]
    method, g, dk, dp, ok, op, n, P, mass, T, x0, spread, dt = case
    ntraj, nsteps, se = 200, 300, 30
    r0, v0 = _inputs(case, ntraj, 5)
    state = (np.arange(ntraj) % n).astype(np.int32)
    print(f"ring-polymer mapping parity: {name}", flush=True)
    dev = _device(L, case, r0, v0, state, nsteps, se, key)
    om = oracle.RPMAP_NRPMD if method == "NRPMD" else oracle.RPMAP_ECMM
    q0, p0 = oracle.rpmap_sample(om, n, P, g, ntraj, key, state=state)
    c = _oracle(oracle, case, r0, v0, q0, p0, nsteps, se)
    _check(dev, c, nsteps)
    assert np.abs(c["pop"][:, -1] - c["pop"][:, 0]).max() > 1e-3   # the populations move


def test_rp_mapping_uploaded_variables(oracle, key):
    """nqcd_mi355x_load_mapping with (ntraj, nbeads, n) variables off the sampling laws."""
    from nqcd_mi355x import _lib as L
    case = CASES["ecmm_tully1_P16"]
    method, g, dk, dp, ok, op, n, P, mass, T, x0, spread, dt = case
    ntraj, nsteps = 100, 200
    r0, v0 = _inputs(case, ntraj, 6)
    rng = np.random.default_rng(8)
    q0, p0 = rng.standard_normal((ntraj, P, n)), rng.standard_normal((ntraj, P, n))
    print("ring-polymer mapping: uploaded variables", flush=True)
    dev = _device(L, case, r0, v0, None, nsteps, 20, key, X=q0, P_=p0)
    c = _oracle(oracle, case, r0, v0, q0, p0, nsteps, 20)
    _check(dev, c, nsteps)


def test_rp_mapping_large_ensemble_invariants(oracle, key):
    """1e5 NRPMD ring polymers of 8 beads on Tully 1 through the crossing: the population sum (1 for every ring
    polymer: the per-bead rotations are orthogonal) to 1e-10, every ring polymer's energy (kinetic + potential +
    springs) to 1e-5 of its kinetic energy, and 128 ring polymers picked across the ensemble against the oracle."""
    from nqcd_mi355x import _lib as L
    case = ("NRPMD", 0.5, "TULLY1", TULLY1, "MODEL_TULLY1", TULLY1, 2, 8, 2000.0, 1e-3, -5.0, 0.05, 1.0)
    ntraj, nsteps, se = 100_000, 800, 100
    rng = np.random.default_rng(12)
    r0 = -5.0 + 0.05 * rng.standard_normal((ntraj, 8))
    v0 = (15.0 + rng.standard_normal((ntraj, 1))) / 2000.0 + 1e-4 * rng.standard_normal((ntraj, 8))
    state = np.zeros(ntraj, np.int32)
    print(f"ring-polymer mapping invariants: ntraj={ntraj} P=8 steps={nsteps}", flush=True)
    dev = _device(L, case, r0, v0, state, nsteps, se, key)
    pop = dev["SAVE_POP_DI"]
    assert np.allclose(pop.sum(axis=2), 1.0, rtol=0, atol=1e-10)
    e = dev["SAVE_EKIN"] + dev["SAVE_EPOT"]
    assert np.all(np.abs(e - e[:, :1]) <= 1e-5 * dev["SAVE_EKIN"][:, :1])
    assert dev["red_pop"][-1, 1] > 0.05
    pick = np.linspace(0, ntraj - 1, 128).astype(np.int64)
    q0, p0 = oracle.rpmap_sample(oracle.RPMAP_NRPMD, 2, 8, 0.5, len(pick), key, state=state[pick], traj_ids=pick)
    c = _oracle(oracle, case, r0[pick], v0[pick], q0, p0, nsteps, se)
    assert np.allclose(pop[pick], c["pop"], rtol=0, atol=1e-9)
    assert np.allclose(dev["R"][pick], c["r"], rtol=1e-9, atol=1e-9)


def test_rp_mapping_run_dynamics(oracle, key):
    """RingPolymerSimulation[NRPMD] through run_dynamics: mean diabatic populations and centroid (device
    reductions) and the per-trajectory OutputFinal against the oracle from the same beads and draws."""
    import nqcd_mi355x as N
    ntraj, P, T, dt, tf = 256, 4, 1e-3, 2.0, 400.0
    sim = N.RingPolymerSimulation[N.NRPMD](N.Atoms(2000.0), N.TullyModelOne(), P, temperature=T, gamma=0.5)
    rng = np.random.default_rng(14)
    r0 = -2.0 + 0.05 * rng.standard_normal((ntraj, P))
    v0 = np.full((ntraj, P), 10.0 / 2000.0)
    dist = N.DynamicalDistribution(list(v0.reshape(ntraj, 1, 1, P)), list(r0.reshape(ntraj, 1, 1, P)), (1, 1, P)) * \
        N.PureState(1, N.Diabatic())
    alg = N.EnsembleMI355X()
    print("ring-polymer mapping run_dynamics", flush=True)
    mean = N.run_dynamics(sim, (0.0, tf), dist, output=(N.OutputDiabaticPopulation, N.OutputCentroidPosition),
                          trajectories=ntraj, dt=dt, saveat=40.0, reduction=N.MeanReduction(), ensemble_algorithm=alg)
    per = N.run_dynamics(sim, (0.0, tf), dist, output=N.OutputFinal, trajectories=ntraj, dt=dt,
                         reduction=N.SortByTrajectoryReduction(), ensemble_algorithm=alg)
    q0, p0 = oracle.rpmap_sample(oracle.RPMAP_NRPMD, 2, P, 0.5, ntraj, alg.seed, state=np.zeros(ntraj, np.int32))
    m = oracle.make_model(oracle.MODEL_TULLY1, 2, 1, TULLY1)
    c = oracle.rpmap_run(oracle.RPMAP_NRPMD, m, 2, P, 2000.0, T, 0.5, r0, v0, q0, p0, dt=dt, nsteps=int(tf / dt),
                         save_every=20)
    assert np.allclose(np.array(mean["OutputDiabaticPopulation"]), c["pop"].mean(axis=0), rtol=0, atol=1e-10)
    assert np.allclose(mean["OutputCentroidPosition"], c["centroid"].mean(axis=0), rtol=1e-10, atol=1e-10)
    assert np.allclose(np.array([p["OutputFinal"]["qmap"] for p in per]), c["q"].transpose(0, 2, 1), rtol=0, atol=1e-9)
    assert np.allclose(np.array([p["OutputFinal"]["r"].ravel() for p in per]), c["r"], rtol=1e-9, atol=1e-9)
