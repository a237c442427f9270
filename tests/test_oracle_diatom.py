"""Oracle diatomic-on-surface model (two Morse diabats + Gaussian coupling, 6 Cartesian DOF):
finite-difference derivatives (as test/Dynamics/utils.jl:39-65 does for every model), lateral
translation / rotation invariance, the Morse limits, and an FSSH energy check on it."""
import numpy as np
import pytest

BASE = [0.24, 1.1, 2.2, 2.0, 1.2, 0.20, 1.0, 2.35, 0.12, 0.8, 3.0, 0.0, 0.003, 1.0, 3.75, 0.5]


def _model(oracle, p=BASE):
    return oracle.make_model(oracle.MODEL_DIATOM_SURFACE, 2, 6, p)


def _config(rng, Z=4.0):
    u = rng.normal(size=3); u /= np.linalg.norm(u)
    b = 2.2 + 0.2 * rng.normal()
    c = np.array([rng.normal(), rng.normal(), Z])
    return np.concatenate([c + 0.5 * b * u, c - 0.5 * b * u])


def test_diatom_derivatives_match_finite_differences(oracle):
    m = _model(oracle)
    rng = np.random.default_rng(0)
    for _ in range(5):
        r = _config(rng, Z=rng.uniform(2.5, 6.0))
        V, dV, _, _ = oracle.model_eval(m, r)
        assert np.allclose(V, V.T)
        h = 1e-6
        for j in range(6):
            rp, rm = r.copy(), r.copy(); rp[j] += h; rm[j] -= h
            fd = (oracle.model_eval(m, rp)[0] - oracle.model_eval(m, rm)[0]) / (2 * h)
            assert np.allclose(dV[j], fd, atol=1e-8, rtol=1e-6)


def test_diatom_invariances_and_limits(oracle):
    m = _model(oracle)
    rng = np.random.default_rng(1)
    r = _config(rng, Z=3.5)
    V, dV, _, _ = oracle.model_eval(m, r)
    shifted = r + np.array([0.7, -1.3, 0.0, 0.7, -1.3, 0.0])          # lateral translation
    assert np.allclose(oracle.model_eval(m, shifted)[0], V, atol=1e-14)
    th = 0.9; R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    rot = np.concatenate([R @ r[:3], R @ r[3:]])                        # rotation about the surface normal
    assert np.allclose(oracle.model_eval(m, rot)[0], V, atol=1e-14)
    assert np.allclose(dV[0] + dV[3], 0, atol=1e-14) and np.allclose(dV[1] + dV[4], 0, atol=1e-14)  # no lateral force
    # far from the surface, at the neutral bond length: V11 -> 0, V22 -> Ds + Di(...) + Delta, V12 -> 0
    far = np.array([0, 0, 50 + 1.1, 0, 0, 50 - 1.1])
    Vf = oracle.model_eval(m, far)[0]
    e = np.exp(-BASE[6] * (2.2 - BASE[7]))
    assert abs(Vf[0, 0]) < 1e-12 and abs(Vf[0, 1]) < 1e-12
    assert np.isclose(Vf[1, 1], BASE[5] * (1 - e) ** 2 + BASE[8] + BASE[11], rtol=1e-10)


def test_diatom_fssh_energy_conservation(oracle, key):
    m = _model(oracle)
    rng = np.random.default_rng(2)
    ntraj = 16
    r0 = np.stack([_config(rng, Z=6.0) for _ in range(ntraj)])
    mass = np.array([12859.0, 12859.0])
    v0 = np.zeros((ntraj, 6)); v0[:, 2] = v0[:, 5] = -3e-3
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    out = oracle.run_ensemble(0, m, np.repeat(mass, 3), r0, v0, sig0, state0=np.zeros(ntraj, np.int32), dt=2.0,
                              nsteps=2500, key=key)
    E = lambda r, v, s: 0.5 * np.sum(np.repeat(mass, 3) * v ** 2) + np.linalg.eigvalsh(oracle.model_eval(m, r)[0])[s]
    for i in range(ntraj):
        e0 = E(r0[i], v0[i], 0)
        e1 = E(out["r"][i], out["v"][i], out["state"][i])
        # velocity Verlet at dt = 2 (bond period ~660 au); a hop adds the stale-force error of one step
        assert abs(e1 - e0) < (1e-4 if out["nhops"][i] == 0 else 5e-3) * abs(e0)
    assert out["nhops"].sum() > 0                       # nonadiabatic: hops at the diabatic crossing
