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


# ---- eCMM (cmm.jl), integrated by the same MInt step ----

@pytest.mark.parametrize("gamma", [0.0, 0.5])
@pytest.mark.parametrize("case", [0, 1, 2])
def test_ecmm_motion_obeys_hamiltons_equations(oracle, key, case, gamma):
    """test/Dynamics/cmm.jl test_motion! at gamma = 0 and 0.5: eCMM's motion! against central differences of its
    Hamiltonian (q^T V q + p^T V p - 2 gamma tr V) / 2 (a gamma term with the wrong factor fails on the
    non-traceless models)."""
    name, m, n, mass, x0 = _models(oracle)[case]
    X, P = oracle.ecmm_sample(n, 1, gamma, key)
    u = np.concatenate([[x0 + 0.3, 2e-3], X[0], P[0]])
    E = lambda w: oracle.mapping_energy(m, n, mass, w, method=oracle.MAP_ECMM, gamma=gamma)
    du = oracle.mapping_motion(m, n, mass, u, method=oracle.MAP_ECMM, gamma=gamma)
    g = np.zeros_like(u)
    for k in range(len(u)):
        h = 1e-6 * max(1.0, abs(u[k]))
        up, um = u.copy(), u.copy()
        up[k] += h; um[k] -= h
        g[k] = (E(up) - E(um)) / (2 * h)
    assert np.allclose(du[0], g[1] / mass, rtol=1e-6)
    assert np.allclose(du[1], -g[0] / mass, rtol=1e-5, atol=1e-12)
    assert np.allclose(du[2:2 + n], g[2 + n:], rtol=1e-6, atol=1e-10)
    assert np.allclose(du[2 + n:], -g[2:2 + n], rtol=1e-6, atol=1e-10)


def test_ecmm_sample_law(oracle, key):
    """cmm.jl:26-43: every sample lies on the sphere of radius sqrt(2 + 2 F gamma) (generate_random_points_on_nsphere,
    test/Dynamics/cmm.jl:56-61), so sum_i (X_i^2 + P_i^2)/2 - gamma (the mapping kernel) is 1; uniform on the sphere,
    each coordinate's square averages R^2 / 2F and the inverse mapping kernel's mean is 1/F per state (5 sigma over
    20000 samples)."""
    n, ntraj, gamma = 3, 20000, 0.3
    X, P = oracle.ecmm_sample(n, ntraj, gamma, key)
    R2 = 2 + 2 * n * gamma
    assert np.allclose((X ** 2).sum(1) + (P ** 2).sum(1), R2, rtol=1e-13)
    assert np.allclose(((X ** 2 + P ** 2) / 2 - gamma).sum(1), 1.0, rtol=1e-13)
    z2 = np.concatenate([X, P], axis=1) ** 2
    assert np.abs(z2.mean(0) - R2 / (2 * n)).max() < 5 * z2.std(0).max() / np.sqrt(ntraj)
    om = oracle.make_model(oracle.MODEL_TULLY1, 2, 1, TULLY1)
    X2, P2 = oracle.ecmm_sample(2, ntraj, gamma, key)
    c = oracle.mapping_run(om, 2, 2000.0, np.full(ntraj, -8.0), np.zeros(ntraj), X2, P2, dt=1.0, nsteps=0,
                           method=oracle.MAP_ECMM, gamma=gamma)
    pop = c["pop"][:, 0]
    assert np.abs(pop.mean(0) - 0.5).max() < 5 * pop.std(0).max() / np.sqrt(ntraj)


def test_ecmm_mint_accuracy_and_gamma_term(oracle, key):
    """MInt on eCMM (cmm.jl, mint.jl:42-90): on Tully 1 (tr dV = 0) at gamma = 0.5 it follows a DOP853 integration
    of eCMM's motion! within 1e-4 and conserves the eCMM energy; on the double well (tr dV != 0) its trace term
    tr(dV) gamma dt / 2m (mint.jl:85) is half of eCMM's, so it follows the motion of gamma / 2 instead and departs
    from that of gamma -- the reference's step, reproduced."""
    gamma, T, dt = 0.5, 300.0, 0.25
    for case, want_half in ((0, False), (1, True)):
        name, m, n, mass, x0 = _models(oracle)[case]
        X, P = oracle.ecmm_sample(n, 1, gamma, key)
        u0 = np.concatenate([[x0, 5e-3], X[0], P[0]])
        run = oracle.mapping_run(m, n, mass, u0[:1], u0[1:2], u0[2:2 + n][None], u0[2 + n:][None], dt=dt,
                                 nsteps=int(T / dt), save_every=int(T / dt) // 10, method=oracle.MAP_ECMM, gamma=gamma)
        b = np.concatenate([run["x"], run["v"], run["X"][0], run["P"][0]])

        def acc(g):
            sol = solve_ivp(lambda t, y: oracle.mapping_motion(m, n, mass, y, method=oracle.MAP_ECMM, gamma=g), (0.0, T),
                            u0, method="DOP853", rtol=1e-12, atol=1e-12)
            return sol.y[:, -1]
        full, half = acc(gamma), acc(gamma / 2)
        scale = np.abs(full).max()
        if want_half:
            assert np.allclose(b, half, rtol=0, atol=1e-4 * scale), np.abs(b - half).max()
            assert np.abs(b - full).max() > 1e-2 * scale
        else:
            assert np.allclose(full, half, rtol=0, atol=1e-9 * scale)   # traceless: the gamma term is absent
            assert np.allclose(b, full, rtol=0, atol=1e-4 * scale), np.abs(b - full).max()
            e = (run["ekin"] + run["epot"])[0]
            assert np.all(np.abs(e - e[0]) <= 1e-4 * abs(e[0]) + 1e-9), e


def test_mapping_energy_closed_form(oracle):
    """The mapping Hamiltonians at chosen points of the double well (V11 = a x^4 - b x^2, V22 = V11 + offset,
    V12 = C exp(-D x^2)): spin mapping sum_i V_ii (X_i^2 + P_i^2 - gamma)/2 + V12 (X1 X2 + P1 P2)
    (spin_mapping.jl:102-118), eCMM (q^T V q + p^T V p - 2 gamma tr V) / 2 (cmm.jl:84-94), each plus m v^2 / 2;
    pins the constant each method subtracts (a factor 2 on gamma fails)."""
    a, b, off, Cc, Dd = DW
    m = oracle.make_model(oracle.MODEL_DW_GAUSSIAN, 2, 1, DW)
    x, v, mass = 1.7, 3e-3, 1836.0
    V11 = a * x ** 4 - b * x ** 2
    V22, V12 = V11 + off, Cc * np.exp(-Dd * x * x)
    X, P = np.array([0.8, -0.3]), np.array([0.4, 1.1])
    quad = (V11 * (X[0] ** 2 + P[0] ** 2) + V22 * (X[1] ** 2 + P[1] ** 2)) / 2 + V12 * (X[0] * X[1] + P[0] * P[1])
    u = np.concatenate([[x, v], X, P])
    gs = oracle.mapping_gamma(2)
    want = 0.5 * mass * v * v + quad - gs / 2 * (V11 + V22)
    assert np.isclose(oracle.mapping_energy(m, 2, mass, u), want, rtol=1e-14, atol=1e-17)
    for g in (0.0, 0.37):
        want = 0.5 * mass * v * v + quad - g * (V11 + V22)
        assert np.isclose(oracle.mapping_energy(m, 2, mass, u, method=oracle.MAP_ECMM, gamma=g), want, rtol=1e-14,
                          atol=1e-17)
