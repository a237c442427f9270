"""GPU two-state FSSH / Ehrenfest on the six-coordinate diatomic-on-surface model against the
FP64 oracle (identical hop sequences, R/P/sigma within 1e-9 over the window) on both schedules (the
persistent kernel and the step-per-launch plain port), and the FP32-nuclei path's final populations
against the FP64 path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = [0.24, 1.1, 2.2, 2.0, 1.2, 0.20, 1.0, 2.35, 0.12, 0.8, 3.0, 0.0, 0.003, 1.0, 3.75, 0.5]
MASS = 12859.0


def _inputs(ntraj, seed, Z=6.0, vz=3e-3):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(ntraj, 3)); u /= np.linalg.norm(u, axis=1, keepdims=True)
    b = 2.2 + 0.1 * rng.normal(size=(ntraj, 1))
    c = np.concatenate([rng.normal(size=(ntraj, 2)), np.full((ntraj, 1), Z)], axis=1)
    r0 = np.concatenate([c + 0.5 * b * u, c - 0.5 * b * u], axis=1)
    v0 = 2e-4 * rng.normal(size=(ntraj, 6))
    v0[:, 2] -= vz; v0[:, 5] -= vz
    return r0, v0


def _gpu(L, method, r0, v0, nsteps, key, precision=0, save_every=50, term=(2, 8.0), schedule=0):
    ntraj = len(r0)
    E = L.Ensemble(0)
    E.set_problem(method, 2, 3, 2, [MASS, MASS], 2.0, precision=precision, schedule=schedule)
    E.set_model(0, L.MODEL_DIATOM_SURFACE, BASE)
    sig = np.zeros((ntraj, 2, 2), complex); sig[:, 0, 0] = 1
    E.load_initial(r0, v0, sig, state=np.zeros(ntraj, np.int32))
    st = E.run(0.0, 2.0 * nsteps, key, save_every=save_every, term=term)
    out = {"r": E.download(L.FIELD_R), "v": E.download(L.FIELD_V), "state": E.download(L.FIELD_STATE),
           "sig": E.download(L.FIELD_SIGMA_RE) + 1j * E.download(L.FIELD_SIGMA_IM), "steps": E.download(L.FIELD_STEPS),
           "nhops": E.download(L.FIELD_NHOPS), "save_state": E.download(L.FIELD_SAVE_STATE),
           "pop_di": E.download(L.FIELD_SAVE_POP_DI), "stats": st,
           "final": E.reduce(L.OUT_FINAL_ADIABATIC_POPULATION, L.RED_MEAN, 2)}
    E.close()
    return out


def _cpu(oracle, method, r0, v0, nsteps, key, save_every=50):
    m = oracle.make_model(oracle.MODEL_DIATOM_SURFACE, 2, 6, BASE)
    ntraj = len(r0)
    sig = np.zeros((ntraj, 2, 2), complex); sig[:, 0, 0] = 1
    return oracle.run_ensemble(method, m, np.full(6, MASS), r0, v0, sig, state0=np.zeros(ntraj, np.int32), dt=2.0,
                               nsteps=nsteps, key=key, save_every=save_every, nsave=nsteps // save_every + 1,
                               term_x=8.0, term_dof=2)


@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("method", [0, 1])
def test_diatom_fp64_parity(oracle, key, method, schedule):
    from nqcd_mi355x import _lib as L
    r0, v0 = _inputs(256, 31)
    nsteps = 1500
    print(f"diatom parity: method {method} schedule {schedule}", flush=True)
    g = _gpu(L, method, r0, v0, nsteps, key, schedule=schedule)
    c = _cpu(oracle, method, r0, v0, nsteps, key)
    assert np.array_equal(g["steps"], c["steps"])
    if method == 0:
        assert c["nhops"].sum() > 20
        assert np.array_equal(g["save_state"], c["save_state"]) and np.array_equal(g["state"], c["state"])
    assert np.allclose(g["r"], c["r"], rtol=1e-9, atol=1e-10)
    assert np.allclose(g["v"], c["v"], rtol=1e-9, atol=1e-13)
    assert np.allclose(g["sig"], c["sig_re"] + 1j * c["sig_im"], rtol=1e-9, atol=1e-10)
    assert np.allclose(g["pop_di"], c["pop_di"], rtol=1e-9, atol=1e-10)


def test_diatom_fp32_population_error(key):
    """FP32 nuclei / FP64 electronics against the FP64 path: final adiabatic populations of 20000
    trajectories within 0.01 (and within 3 sigma of the difference of two ensemble means)."""
    from nqcd_mi355x import _lib as L
    r0, v0 = _inputs(20000, 32)
    g64 = _gpu(L, 0, r0, v0, 5000, key, precision=0, save_every=0)
    g32 = _gpu(L, 0, r0, v0, 5000, key, precision=1, save_every=0)
    p64, p32 = g64["final"][1], g32["final"][1]
    sigma = np.sqrt(2 * p64 * (1 - p64) / len(r0))
    assert abs(p64 - p32) < 0.01 and abs(p64 - p32) < 3 * sigma + 1e-12, (p64, p32, sigma)
    assert 0.05 < p64 < 0.95


def test_diatom_autotune_keeps_trajectories(key):
    """The autotuner on the diatom-surface kernel (family 2): the tuned run reproduces the default
    run's trajectories bit for bit."""
    from nqcd_mi355x import _lib as L
    r0, v0 = _inputs(2000, 33)
    ref = _gpu(L, 0, r0, v0, 400, key, precision=1)
    E = L.Ensemble(0)
    E.set_problem(0, 2, 3, 2, [MASS, MASS], 2.0, precision=1)
    E.set_model(0, L.MODEL_DIATOM_SURFACE, BASE)
    sig = np.zeros((len(r0), 2, 2), complex); sig[:, 0, 0] = 1
    E.load_initial(r0, v0, sig, state=np.zeros(len(r0), np.int32))
    res = E.autotune(0.0, 100.0, key)
    E.close()
    assert res["family"] == 2 and res["n_candidates"] >= 6 and res["block"] in (64, 128, 256)
    tuned = _gpu(L, 0, r0, v0, 400, key, precision=1)
    assert tuned["stats"]["block"] == res["block"]
    for k in ("r", "v", "state", "steps", "nhops", "save_state"):
        assert np.array_equal(tuned[k], ref[k]), k


def test_diatom_plain_schedule_rejects_fp32_and_surrogate(key):
    """The step-per-launch port runs the analytic model in FP64 only: FP32 nuclei are refused with a
    clear error before any launch."""
    from nqcd_mi355x import _lib as L
    r0, v0 = _inputs(64, 35)
    with pytest.raises(L.NQCDError, match="FP64"):
        _gpu(L, 0, r0, v0, 10, key, precision=1, schedule=1)
