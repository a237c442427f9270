"""Cell callbacks on the device (CellBoundaryCallback / TerminateCellCallback, DynamicsUtils/callbacks.jl:4-22,
on NQCBase's PeriodicCell) and classical VelocityVerlet dynamics (Simulation{Classical}, classical.jl:80-86)
against the FP64 oracle: identical hop sequences, exit steps and wraps, positions to 1e-9; and the
reference's own cell test (test/Dynamics/cell_boundary_callback.jl) through run_dynamics."""
import numpy as np
import pytest

import nstate_models as NS

pytestmark = pytest.mark.gpu

TULLY1 = [0.01, 1.6, 0.005, 1.0]
HEX = np.array([[5.883, -2.942, 0.0], [0.0, 5.095, 0.0], [0.0, 0.0, 20.0]])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_classical_velocity_verlet_matches_oracle(oracle, mode):
    from nqcd_mi355x import _lib as L
    ntraj, nsteps = 1000, 2000
    rng = np.random.default_rng(7)
    r0 = rng.uniform(0.1, 1.9, ntraj)
    v0 = rng.normal(0.0, 0.02, ntraj)
    prm = [1.0, 0.05, 1.0]                       # V = 0.0025 (x - 1)^2 / 2
    print(f"classical VelocityVerlet: cell mode {mode}", flush=True)
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_CLASSICAL, 1, 1, 1, [1.0], 0.5, algorithm=L.ALG_VELOCITY_VERLET)
    E.set_model(0, L.MODEL_HARMONIC, prm)
    if mode:
        E.set_cell(mode, [[2.0]])
    E.load_initial(r0[:, None], v0[:, None], None)
    E.run(0.0, 0.5 * nsteps, 0, save_every=10)
    g = {f: E.download(getattr(L, "FIELD_" + f)) for f in ("R", "V", "STEPS", "SAVE_R", "SAVE_V", "SAVE_EKIN", "SAVE_EPOT")}
    red = {k: E.reduce(getattr(L, "OUT_" + k), L.RED_MEAN, E.nsave) for k in ("POSITION", "VELOCITY", "TOTAL_ENERGY")}
    E.close()
    m = oracle.make_model(oracle.MODEL_HARMONIC, 1, 1, prm)
    cell = oracle.make_cell(mode, [[2.0]]) if mode else None
    c = oracle.classical_run(m, 1.0, r0, v0, dt=0.5, nsteps=nsteps, save_every=10, cell=cell)
    assert np.array_equal(g["STEPS"], c["steps"])
    if mode == 2:
        assert np.any(c["steps"] < nsteps)
    assert np.allclose(g["R"][:, 0], c["x"], rtol=1e-10, atol=1e-12)
    assert np.allclose(g["V"][:, 0], c["v"], rtol=1e-10, atol=1e-14)
    assert np.allclose(g["SAVE_R"], c["save_x"], rtol=1e-10, atol=1e-12)
    assert np.allclose(g["SAVE_V"], c["save_v"], rtol=1e-10, atol=1e-14)
    assert np.allclose(g["SAVE_EKIN"] + g["SAVE_EPOT"], c["ekin"] + c["epot"], rtol=1e-10, atol=1e-15)
    assert np.allclose(red["POSITION"], c["save_x"].mean(axis=0), rtol=1e-10, atol=1e-12)
    assert np.allclose(red["VELOCITY"], c["save_v"].mean(axis=0), rtol=1e-10, atol=1e-14)
    assert np.allclose(red["TOTAL_ENERGY"], (c["ekin"] + c["epot"]).mean(axis=0), rtol=1e-10, atol=1e-15)


def test_reference_cell_boundary_test_through_run_dynamics():
    """test/Dynamics/cell_boundary_callback.jl: Atoms([:C]), Harmonic(), PeriodicCell(hcat(1)), v = 1, r = 0,
    run_dynamics over (0, 300) with dt = 0.1 and OutputPosition: the atom leaves the cell without the
    callback and stays in it with CellBoundaryCallback."""
    import nqcd_mi355x as N
    sim = N.Simulation[N.Classical](N.Atoms("C"), N.Harmonic(), cell=N.PeriodicCell([[1.0]]))
    z = N.DynamicsVariables(sim, np.ones((1, 1)), np.zeros((1, 1)))
    sol = N.run_dynamics(sim, (0.0, 300.0), z, output=N.OutputPosition, dt=0.1)
    Rs = np.array([float(np.ravel(x)[0]) for x in sol["OutputPosition"]])
    assert len(Rs) == 3001 and Rs.max() > 1.0
    sol = N.run_dynamics(sim, (0.0, 300.0), z, output=N.OutputPosition, callback=N.CellBoundaryCallback(), dt=0.1)
    Rs = np.array([float(np.ravel(x)[0]) for x in sol["OutputPosition"]])
    assert np.isclose(Rs.min(), 0, atol=1e-4) or Rs.min() >= 0


def _tully(L, oracle, key, mode, schedule, precision=0, ntraj=512, nsteps=1500):
    rng = np.random.default_rng(3)
    r0 = rng.uniform(1.0, 7.0, size=(ntraj, 1))
    v0 = rng.uniform(-30, 30, size=(ntraj, 1)) / 2000.0
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_FSSH, 2, 1, 1, [2000.0], 1.0, schedule=schedule, precision=precision)
    E.set_model(0, L.MODEL_TULLY1, TULLY1)
    E.set_cell(mode, [[8.0]])
    E.load_initial(r0, v0, sig0, state=np.zeros(ntraj, np.int32))
    E.run(0.0, float(nsteps), key, save_every=50)
    g = {f: E.download(getattr(L, "FIELD_" + f)) for f in ("R", "V", "STATE", "NHOPS", "STEPS", "SAVE_STATE")}
    E.close()
    m = oracle.make_model(oracle.MODEL_TULLY1, 2, 1, TULLY1)
    c = oracle.run_ensemble(0, m, 2000.0, r0, v0, sig0, state0=np.zeros(ntraj, np.int32), dt=1.0, nsteps=nsteps,
                            key=key, save_every=50, nsave=nsteps // 50 + 1, cell=oracle.make_cell(mode, [[8.0]]))
    return g, c


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("schedule", [0, 1])
def test_fssh_cell_callbacks_match_oracle(oracle, key, mode, schedule):
    from nqcd_mi355x import _lib as L
    print(f"FSSH cell callbacks: mode {mode} schedule {schedule}", flush=True)
    g, c = _tully(L, oracle, key, mode, schedule)
    assert c["nhops"].sum() > 0
    if mode == 2:
        assert np.any(c["steps"] < 1500)
    else:
        assert np.all((c["r"] >= 0) & (c["r"] <= 8))
    assert np.array_equal(g["STEPS"], c["steps"])
    assert np.array_equal(g["STATE"], c["state"]) and np.array_equal(g["NHOPS"], c["nhops"])
    assert np.array_equal(g["SAVE_STATE"], c["save_state"])
    assert np.allclose(g["R"], c["r"], rtol=1e-9, atol=1e-10)
    assert np.allclose(g["V"], c["v"], rtol=1e-9, atol=1e-13)


def test_fssh_fp32_cell_wrap_keeps_atoms_in_the_cell(oracle, key):
    from nqcd_mi355x import _lib as L
    g, c = _tully(L, oracle, key, 1, 0, precision=L.PREC_FP32_NUCLEI, ntraj=2048)
    assert np.all((g["R"] >= 0) & (g["R"] <= 8))
    pg = np.bincount(g["STATE"], minlength=2) / len(g["STATE"])
    pc = np.bincount(c["state"], minlength=2) / len(c["state"])
    assert np.all(np.abs(pg - pc) < 5 * np.sqrt(pc * (1 - pc) / len(c["state"])) + 0.02), (pg, pc)


@pytest.mark.parametrize("mode", [1, 2])
def test_nstate_cell_callbacks_match_oracle(oracle, key, mode):
    from nqcd_mi355x import _lib as L
    n, ntraj, nsteps = 3, 256, 1200
    rng = np.random.default_rng(9)
    r0 = rng.normal(3.0, 0.15, size=(ntraj, 1))
    v0 = np.sqrt(2 * rng.uniform(0.02, 0.04, size=(ntraj, 1)) / NS.MASS)
    sig0 = np.zeros((ntraj, n, n), complex); sig0[:, 0, 0] = 1
    lo, width = 2.6, 1.4                                   # cell [2.6, 4.0) along x: a shifted 1 x 1 cell
    print(f"n-state cell callbacks: mode {mode}", flush=True)
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_FSSH, n, 1, 1, [NS.MASS], 1.0)
    E.set_model(0, L.MODEL_MORSE_N, [n] + NS.three_state())
    E.set_cell(mode, [[width]])
    E.load_initial(r0, v0, sig0, state=np.zeros(ntraj, np.int32))
    E.run(0.0, float(nsteps), key)
    g = {f: E.download(getattr(L, "FIELD_" + f)) for f in ("R", "STATE", "NHOPS", "STEPS")}
    E.close()
    m = oracle.make_model(oracle.MODEL_MORSE_N, n, 1, NS.three_state())
    c = oracle.run_ensemble(0, m, NS.MASS, r0, v0, sig0, state0=np.zeros(ntraj, np.int32), dt=1.0, nsteps=nsteps,
                            key=key, cell=oracle.make_cell(mode, [[width]]))
    assert lo > 0
    assert np.array_equal(g["STEPS"], c["steps"]) and np.array_equal(g["STATE"], c["state"])
    assert np.array_equal(g["NHOPS"], c["nhops"])
    assert np.allclose(g["R"], c["r"], rtol=1e-9, atol=1e-10)


def test_cme_terminate_cell(oracle, key):
    from nqcd_mi355x import _lib as L
    hw, Gam, kT, g_, Ed = 0.003, 0.003, 0.03, 0.005, 0.0
    params = [hw, Ed, np.sqrt(2) * g_, np.sqrt(Gam / (2 * np.pi))]
    mass, ntraj, nsteps = 1 / hw, 512, 3000
    prm = [2.0, 1.0, 0.0, np.sqrt(kT / mass)]
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_CME, 2, 1, 1, [mass], 5.0, algorithm=L.ALG_TSIT5, rescaling=L.RESCALE_OFF, temperature=kT)
    E.set_model(0, L.MODEL_ANDERSON_HOLSTEIN_1, params)
    E.set_cell(L.CELL_TERMINATE, [[4.0]])
    E.sample_initial(L.SAMPLER_GAUSSIAN, ntraj, prm, key)
    E.run(0.0, 5.0 * nsteps, key)
    g = {f: E.download(getattr(L, "FIELD_" + f)) for f in ("R", "STATE", "STEPS")}
    E.close()
    r0, v0 = oracle.sample_gaussian(ntraj, 1, *prm, key)
    m = oracle.make_model(oracle.MODEL_ANDERSON_HOLSTEIN_1, 2, 1, params)
    c = oracle.cme_run(0, m, mass, kT, r0[:, 0], v0[:, 0], np.zeros(ntraj, np.int32), dt=5.0, nsteps=nsteps, key=key,
                       cell=oracle.make_cell(2, [[4.0]]))
    assert np.any(c["x"] < 0) and np.any(c["x"] > 4)
    assert np.array_equal(g["STATE"], c["state"])
    assert np.any(g["STEPS"] < nsteps)
    assert np.allclose(g["R"][:, 0], c["x"], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("mode", [1, 2])
def test_diatom_cell_callbacks_match_oracle(oracle, key, mode, schedule):
    """Two atoms of three coordinates in a hexagonal cell (docs/src/initialconditions/ebk.md:47), periodic in
    x and y; the molecule starts near a cell edge so atoms cross it."""
    from nqcd_mi355x import _lib as L
    from test_gpu_diatom import BASE, MASS, _inputs
    r0, v0 = _inputs(256, 41, Z=6.0)
    r0[:, 0] += 0.3; r0[:, 3] += 0.3
    r0[:, 1] += 2.5; r0[:, 4] += 2.5
    nsteps = 1200
    print(f"diatom cell callbacks: mode {mode} schedule {schedule}", flush=True)
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_FSSH, 2, 3, 2, [MASS, MASS], 2.0, schedule=schedule)
    E.set_model(0, L.MODEL_DIATOM_SURFACE, BASE)
    E.set_cell(mode, HEX, [1, 1, 0])
    sig = np.zeros((len(r0), 2, 2), complex); sig[:, 0, 0] = 1
    E.load_initial(r0, v0, sig, state=np.zeros(len(r0), np.int32))
    E.run(0.0, 2.0 * nsteps, key)
    g = {f: E.download(getattr(L, "FIELD_" + f)) for f in ("R", "V", "STATE", "STEPS")}
    E.close()
    m = oracle.make_model(oracle.MODEL_DIATOM_SURFACE, 2, 6, BASE)
    c = oracle.run_ensemble(0, m, np.full(6, MASS), r0, v0, sig, state0=np.zeros(len(r0), np.int32), dt=2.0,
                            nsteps=nsteps, key=key, cell=oracle.make_cell(mode, HEX, [1, 1, 0]))
    if mode == 2:
        assert np.any(c["steps"] < nsteps)
    assert np.array_equal(g["STEPS"], c["steps"]) and np.array_equal(g["STATE"], c["state"])
    assert np.allclose(g["R"], c["r"], rtol=1e-9, atol=1e-9)
    assert np.allclose(g["V"], c["v"], rtol=1e-9, atol=1e-12)
