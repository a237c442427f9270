"""Pins the oracle's ring-polymer mapping dynamics (NRPMD, nrpmd.jl; ring-polymer eCMM, rpecmm.jl) and its
RingPolymerMInt integrator (ringpolymer_mint.jl:28-131) the way the reference's tests do (test/Dynamics/nrpmd.jl,
rpecmm.jl): motion! against finite differences of the ring-polymer Hamiltonian, energy conservation,
RingPolymerMInt against an accurate integration of motion!, second-order convergence; plus the free ring
polymer against the full Cayley step, and the initial-condition laws."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

DW = [1 / 1200, 1 / 40, 0.01, 0.005, 1.0]
TULLY2 = [0.1, 0.28, 0.015, 0.06, 0.05]
MORSE3 = [0.02, 0.5, 4.0, 0.0, 0.02, 0.5, 4.0, 0.004, 0.02, 0.5, 4.0, 0.008,
          0.003, 0.5, 4.0, 0.002, 0.5, 4.2, 0.003, 0.5, 4.4]


def _cases(O):
    # (method, gamma, model, n, P, mass, temperature, x0)
    return [(O.RPMAP_NRPMD, 0.5, O.make_model(O.MODEL_DW_GAUSSIAN, 2, 1, DW), 2, 4, 1836.0, 1 / 8, -3.0),
            (O.RPMAP_ECMM, 0.0, O.make_model(O.MODEL_TULLY2, 2, 1, TULLY2), 2, 4, 2000.0, 1e-3, -1.0),
            (O.RPMAP_ECMM, 0.5, O.make_model(O.MODEL_DW_GAUSSIAN, 2, 1, DW), 2, 3, 1836.0, 1 / 8, -3.0),
            (O.RPMAP_NRPMD, 0.3, O.make_model(O.MODEL_MORSE_N, 3, 1, MORSE3), 3, 4, 20000.0, 1e-3, 4.1)]


def _u0(O, case, key, traj=0, spread=0.05):
    method, g, m, n, P, mass, T, x0 = case
    q, p = O.rpmap_sample(method, n, P, g, 1, key, state=np.array([0]), traj_ids=np.array([traj]))
    rng = np.random.default_rng(traj + 3)
    r = x0 + spread * rng.standard_normal(P)
    v = 2e-3 + 1e-4 * rng.standard_normal(P)
    return np.concatenate([r, v, q.ravel(), p.ravel()])


@pytest.mark.parametrize("c", range(4))
def test_rp_mapping_motion_obeys_hamiltons_equations(oracle, key, c):
    """test/Dynamics/nrpmd.jl "motion! obeys Hamilton's equations" (and rpecmm.jl:12-34): dr = dH/dv / m,
    dv = -dH/dr / m, dq = dH/dp, dp = -dH/dq by central differences of classical_hamiltonian with the springs."""
    case = _cases(oracle)[c]
    method, g, m, n, P, mass, T, x0 = case
    u = _u0(oracle, case, key)
    du = oracle.rpmap_motion(method, m, n, P, mass, T, g, u)
    E = lambda w: oracle.rpmap_energy(method, m, n, P, mass, T, g, w)
    grad = np.zeros_like(u)
    for k in range(len(u)):
        h = 1e-6 * max(1.0, abs(u[k]))
        up, um = u.copy(), u.copy()
        up[k] += h; um[k] -= h
        grad[k] = (E(up) - E(um)) / (2 * h)
    # central differences of H (springs of size ~1 at these temperatures) carry ~1e-10 of rounding
    assert np.allclose(du[:P], grad[P:2 * P] / mass, rtol=1e-6)
    assert np.allclose(du[P:2 * P], -grad[:P] / mass, rtol=1e-5, atol=1e-6 * np.abs(du[P:2 * P]).max())
    nq = P * n
    tol = 1e-6 * np.abs(du[2 * P:]).max()
    assert np.allclose(du[2 * P:2 * P + nq], grad[2 * P + nq:], rtol=1e-6, atol=tol)
    assert np.allclose(du[2 * P + nq:], -grad[2 * P:2 * P + nq], rtol=1e-6, atol=tol)


def _mint(O, case, u0, dt, nsteps, save_every=0):
    method, g, m, n, P, mass, T, x0 = case
    nq = P * n
    return O.rpmap_run(method, m, n, P, mass, T, g, u0[:P], u0[P:2 * P], u0[2 * P:2 * P + nq], u0[2 * P + nq:], dt=dt,
                       nsteps=nsteps, save_every=save_every)


def _flat(out):
    return np.concatenate([out["r"].ravel(), out["v"].ravel(), out["q"].ravel(), out["p"].ravel()])


def _accurate(O, case, u0, T_end):
    method, g, m, n, P, mass, T, x0 = case
    sol = solve_ivp(lambda t, y: O.rpmap_motion(method, m, n, P, mass, T, g, y), (0.0, T_end), u0, method="DOP853",
                    rtol=1e-12, atol=1e-12)
    return sol.y[:, -1]


@pytest.mark.parametrize("c", range(4))
def test_rp_mint_energy_and_accuracy(oracle, key, c):
    """nrpmd.jl "Energy conservation" (rtol 1e-2) and "Algorithm comparison" (RingPolymerMInt against Tsit5 at
    1e-10 within 1e-2): here the energy of every save within 1e-4 of its start and the state after 200 steps within
    1e-4 of DOP853 at 1e-12."""
    case = _cases(oracle)[c]
    u0 = _u0(oracle, case, key, traj=1)
    # the Cayley step turns a mode of frequency w by 2 atan(w dt / 2), accurate for w dt << 1: dt keeps the
    # highest ring-polymer frequency 2 P T below 0.05 / dt
    P, T = case[4], case[6]
    dt, nsteps = min(2.0, 0.05 / (2 * P * T)), 200
    out = _mint(oracle, case, u0, dt, nsteps, save_every=20)
    e = (out["ekin"] + out["epot"])[0]
    assert np.all(np.abs(e - e[0]) <= 1e-4 * abs(e[0]) + 1e-7), e
    a = _accurate(oracle, case, u0, dt * nsteps)
    b = _flat(out)
    assert np.allclose(b, a, rtol=0, atol=1e-4 * np.abs(a).max()), np.abs(b - a).max()


def test_rp_mint_second_order(oracle, key):
    """nrpmd.jl "MInt algorithm convergence": the final-state error falls as dt^2 (order 2 within 0.15)."""
    case = _cases(oracle)[3]
    u0 = _u0(oracle, case, key, traj=2)
    T_end = 128.0
    ref = _accurate(oracle, case, u0, T_end)
    dts = [4.0, 2.0, 1.0, 0.5]
    err = [np.abs(_flat(_mint(oracle, case, u0, dt, int(round(T_end / dt)))) - ref).max() for dt in dts]
    order = np.polyfit(np.log(dts), np.log(err), 1)[0]
    assert abs(order - 2.0) < 0.15, (order, err)


def test_rp_mint_free_ring_polymer(oracle, key):
    """With a constant model (no force) a step is two half steps sqrt(cay(dt A)) (ring_polymer.jl:71-82, half =
    true), i.e. one full Cayley step per normal mode: n steps equal the full Cayley map applied n times
    (oracle_rp_normal_modes), and the mapping variables only rotate."""
    P, dt, nsteps, T, mass = 8, 3.0, 50, 0.01, 1836.0
    m = oracle.make_model(oracle.MODEL_DW_GAUSSIAN, 2, 1, [0.0, 0.0, 0.01, 0.0, 1.0])
    case = (oracle.RPMAP_NRPMD, 0.5, m, 2, P, mass, T, 0.0)
    u0 = _u0(oracle, case, key, traj=4, spread=0.3)
    out = _mint(oracle, case, u0, dt, nsteps)
    U, w, Cf = oracle.rp_normal_modes(P, P * T, dt)
    rn, vn = U.T @ u0[:P], U.T @ u0[P:2 * P]
    for _ in range(nsteps):
        C = Cf.reshape(P, 2, 2)
        rn, vn = C[:, 0, 0] * rn + C[:, 0, 1] * vn, C[:, 1, 0] * rn + C[:, 1, 1] * vn
    assert np.allclose(out["r"][0], U @ rn, rtol=0, atol=1e-12)
    assert np.allclose(out["v"][0], U @ vn, rtol=0, atol=1e-14)
    pop0 = (u0[2 * P:2 * P + 2 * P] ** 2 + u0[4 * P:] ** 2)
    assert np.allclose(out["q"].ravel() ** 2 + out["p"].ravel() ** 2, pop0, rtol=1e-12)


def test_rp_mapping_initial_laws(oracle, key):
    """nrpmd.jl:49-65: radius sqrt(2 + 2 gamma) on the occupied state and sqrt(2 gamma) elsewhere on every bead, so
    the NRPMD estimator (bead mean of (q^2 + p^2)/2 - gamma) reads exactly 1 and 0; phases uniform (5 sigma).
    rpecmm.jl:11-25: a point on the sphere of radius sqrt(2 + 2 F gamma N) in 2 F N dimensions; the bead sum of the
    inverse mapping kernel has mean 1/F per state (5 sigma)."""
    n, P, g, ntraj = 3, 4, 0.4, 20000
    st = np.arange(ntraj) % n
    q, p = oracle.rpmap_sample(oracle.RPMAP_NRPMD, n, P, g, ntraj, key, state=st)
    r2 = q ** 2 + p ** 2
    want = np.where(np.arange(n)[None, None, :] == st[:, None, None], 2 + 2 * g, 2 * g)
    assert np.allclose(r2, np.broadcast_to(want, r2.shape), rtol=1e-14)
    c = q / np.sqrt(r2)
    assert np.abs(c.mean(0)).max() < 5 / np.sqrt(ntraj)
    m = oracle.make_model(oracle.MODEL_MORSE_N, 3, 1, MORSE3)
    out = oracle.rpmap_run(oracle.RPMAP_NRPMD, m, n, P, 20000.0, 1e-3, g, np.full((ntraj, P), 4.0), np.zeros((ntraj, P)),
                           q, p, dt=1.0, nsteps=0)
    assert np.allclose(out["pop"][:, 0], np.eye(n)[st], atol=1e-14)
    qe, pe = oracle.rpmap_sample(oracle.RPMAP_ECMM, n, P, g, ntraj, key)
    R2 = 2 + 2 * n * g * P
    assert np.allclose((qe ** 2).sum((1, 2)) + (pe ** 2).sum((1, 2)), R2, rtol=1e-13)
    out = oracle.rpmap_run(oracle.RPMAP_ECMM, m, n, P, 20000.0, 1e-3, g, np.full((ntraj, P), 4.0), np.zeros((ntraj, P)),
                           qe, pe, dt=1.0, nsteps=0)
    pop = out["pop"][:, 0]
    assert np.abs(pop.mean(0) - 1 / n).max() < 5 * pop.std(0).max() / np.sqrt(ntraj)


def test_mapping_initial_estimators(oracle, key):
    """The initial estimators of the population correlation: NRPMD's is its population (Estimators.jl:121), so for
    PureState(1) the correlation at t0 is [[1, 0], [0, 0]] exactly (test/Dynamics/nrpmd.jl "Population
    correlation", which allows 0.1); eCMM's mapping kernel (X^2 + P^2)/2 - gamma summed over the states (and beads)
    is 1 for every sample (cmm.jl:106-110, rpecmm.jl:95-103: the sphere radius)."""
    n, P, ntraj = 2, 4, 500
    m = oracle.make_model(oracle.MODEL_TULLY1, 2, 1, [0.01, 1.6, 0.005, 1.0])
    r0, v0 = np.full((ntraj, P), -8.0), np.zeros((ntraj, P))
    q, p = oracle.rpmap_sample(oracle.RPMAP_NRPMD, n, P, 0.5, ntraj, key, state=np.zeros(ntraj, np.int32))
    c = oracle.rpmap_run(oracle.RPMAP_NRPMD, m, n, P, 2000.0, 1e-3, 0.5, r0, v0, q, p, dt=1.0, nsteps=0)
    corr = np.einsum("ta,tb->ab", c["init"], c["pop"][:, 0]) / ntraj
    assert np.allclose(corr, [[1, 0], [0, 0]], atol=1e-13)
    for g in (0.0, 0.3):
        qe, pe = oracle.rpmap_sample(oracle.RPMAP_ECMM, n, P, g, ntraj, key)
        ce = oracle.rpmap_run(oracle.RPMAP_ECMM, m, n, P, 2000.0, 1e-3, g, r0, v0, qe, pe, dt=1.0, nsteps=0)
        assert np.allclose(ce["init"].sum(1), 1.0, rtol=1e-13)
        X, Pm = oracle.ecmm_sample(n, ntraj, g, key)
        cm = oracle.mapping_run(m, n, 2000.0, r0[:, 0], v0[:, 0], X, Pm, dt=1.0, nsteps=0, method=oracle.MAP_ECMM, gamma=g)
        assert np.allclose(cm["init"].sum(1), 1.0, rtol=1e-13)
