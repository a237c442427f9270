"""Pins the oracle's spin-mapping dynamics (SpinMappingW, spin_mapping.jl) and its MInt integrator
(mint.jl:42-181) the way the reference's own tests do (test/Dynamics/spin_mapping.jl): motion! against
finite differences of the Hamiltonian, energy conservation, MInt against a high-accuracy integration of
motion!, and second-order convergence; plus the initial-condition law of spin_mapping.jl:18-37."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

TULLY1 = [0.01, 1.6, 0.005, 1.0]
DW = [1 / 1200, 1 / 40, 0.01, 0.005, 1.0]
MORSE3 = [0.02, 0.4, 4.0, 0.0, 0.02, 0.65, 4.5, 0.01, 0.003, 0.65, 6.0, 0.0,
          0.005, 0.3, 3.5, 0.005, 0.3, 4.5, 0.002, 0.3, 5.0]


def _models(O):
    return [("tully1", O.make_model(O.MODEL_TULLY1, 2, 1, TULLY1), 2, 2000.0, -1.0),
            ("double_well", O.make_model(O.MODEL_DW_GAUSSIAN, 2, 1, DW), 2, 1836.0, -0.5),
            ("morse3", O.make_model(O.MODEL_MORSE_N, 3, 1, MORSE3), 3, 20000.0, 4.3)]


def _u0(O, n, x0, v0, state, key, traj=0):
    X, P = O.mapping_sample(n, np.array([state]), key, traj_ids=np.array([traj]))
    return np.concatenate([[x0, v0], X[0], P[0]])


@pytest.mark.parametrize("case", [0, 1, 2])
def test_motion_obeys_hamiltons_equations(oracle, key, case):
    """spin_mapping.jl test 1: dr = dH/dv / m, dv = -dH/dr / m, dX = dH/dP, dP = -dH/dX, the gradient of H
    by central differences."""
    name, m, n, mass, x0 = _models(oracle)[case]
    u = _u0(oracle, n, x0 + 0.3, 2e-3, 0, key)
    du = oracle.mapping_motion(m, n, mass, u)
    g = np.zeros_like(u)
    for k in range(len(u)):
        h = 1e-6 * max(1.0, abs(u[k]))
        up, um = u.copy(), u.copy()
        up[k] += h; um[k] -= h
        g[k] = (oracle.mapping_energy(m, n, mass, up) - oracle.mapping_energy(m, n, mass, um)) / (2 * h)
    assert np.allclose(du[0], g[1] / mass, rtol=1e-6)
    assert np.allclose(du[1], -g[0] / mass, rtol=1e-5, atol=1e-12)
    assert np.allclose(du[2:2 + n], g[2 + n:], rtol=1e-6, atol=1e-10)
    assert np.allclose(du[2 + n:], -g[2:2 + n], rtol=1e-6, atol=1e-10)


def _accurate(oracle, m, n, mass, u0, T):
    sol = solve_ivp(lambda t, y: oracle.mapping_motion(m, n, mass, y), (0.0, T), u0, method="DOP853", rtol=1e-12,
                    atol=1e-12)
    return sol.y[:, -1]


def _mint(oracle, m, n, mass, u0, dt, nsteps):
    out = oracle.mapping_run(m, n, mass, u0[:1], u0[1:2], u0[2:2 + n][None], u0[2 + n:][None], dt=dt, nsteps=nsteps)
    return np.concatenate([out["x"], out["v"], out["X"][0], out["P"][0]])


@pytest.mark.parametrize("case", [0, 1, 2])
def test_mint_energy_conservation(oracle, key, case):
    """spin_mapping.jl test 2 (OutputTotalEnergy at the first and last save within 1e-2): here every save within
    1e-4 relative over 2000 steps, with kinetic + potential energy from the per-save outputs."""
    name, m, n, mass, x0 = _models(oracle)[case]
    u0 = _u0(oracle, n, x0, 5e-3 if n == 2 else 1e-3, 0, key)
    out = oracle.mapping_run(m, n, mass, u0[:1], u0[1:2], u0[2:2 + n][None], u0[2 + n:][None], dt=0.5, nsteps=2000,
                             save_every=100)
    e = (out["ekin"] + out["epot"])[0]
    assert np.all(np.abs(e - e[0]) <= 1e-4 * abs(e[0]) + 1e-9), e


@pytest.mark.parametrize("case", [0, 1, 2])
def test_mint_matches_accurate_integration(oracle, key, case):
    """spin_mapping.jl test 3 (MInt at dt against Tsit5 at 1e-10, within 1e-2): here against DOP853 at 1e-12,
    within 1e-4 of the largest component."""
    name, m, n, mass, x0 = _models(oracle)[case]
    u0 = _u0(oracle, n, x0, 5e-3 if n == 2 else 1e-3, 1, key)
    T, dt = 300.0, 0.25
    a = _accurate(oracle, m, n, mass, u0, T)
    b = _mint(oracle, m, n, mass, u0, dt, int(round(T / dt)))
    scale = np.abs(a).max()
    assert np.allclose(b, a, rtol=0, atol=1e-4 * scale), np.abs(b - a).max()


def test_mint_second_order(oracle, key):
    """spin_mapping.jl test 4: the final-state error of MInt falls as dt^2 (order 2 within 0.1, over dt halving
    from 8 to 0.5 on the single avoided crossing)."""
    name, m, n, mass, x0 = _models(oracle)[0]
    u0 = _u0(oracle, n, -0.8, 8e-3, 0, key)
    T = 256.0
    ref = _accurate(oracle, m, n, mass, u0, T)
    dts = [8.0, 4.0, 2.0, 1.0, 0.5]
    err = [np.abs(_mint(oracle, m, n, mass, u0, dt, int(round(T / dt))) - ref).max() for dt in dts]
    order = np.polyfit(np.log(dts), np.log(err), 1)[0]
    assert abs(order - 2.0) < 0.1, (order, err)


def test_mapping_initial_conditions(oracle, key):
    """spin_mapping.jl:18-37: X_i^2 + P_i^2 = 2 + gamma on the occupied state and gamma elsewhere, so the
    population estimator (X^2 + P^2 - gamma) / 2 reads 1 and 0; the phases are uniform (first and second
    moments of sin and cos over 20000 draws within 5 sigma)."""
    n, ntraj = 3, 20000
    g = oracle.mapping_gamma(n)
    assert np.isclose(g, 2 / 3 * (2 - 1))
    state = np.arange(ntraj) % n
    X, P = oracle.mapping_sample(n, state, key)
    r2 = X ** 2 + P ** 2
    want = np.where(np.arange(n)[None, :] == state[:, None], 2 + g, g)
    assert np.allclose(r2, want, rtol=1e-14)
    pop = (r2 - g) / 2
    assert np.allclose(pop, np.eye(n)[state], atol=1e-14)
    s, c = X / np.sqrt(r2), P / np.sqrt(r2)
    tol = 5 / np.sqrt(ntraj)
    assert np.abs(s.mean(0)).max() < tol and np.abs(c.mean(0)).max() < tol
    assert np.abs((s ** 2).mean(0) - 0.5).max() < 5 * np.sqrt(0.125 / ntraj)
