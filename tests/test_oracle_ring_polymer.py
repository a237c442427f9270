"""Pins the ring-polymer part of the oracle: normal modes and Cayley propagator
(test/Core/ring_polymers.jl:13,25-32), BCB second-order convergence against the
analytic harmonic solution (test/Dynamics/algorithms/bcb.jl:18-44), energy
conservation, and the one-bead limit of RPSH equal to FSSH."""
import numpy as np
import pytest


def spring_matrix(P, wn):  # ring_polymer.jl:46-57
    if P == 1:
        S = np.zeros((1, 1))
    elif P == 2:
        S = np.array([[2.0, -2.0], [-2.0, 2.0]])
    else:
        S = 2 * np.eye(P) - np.eye(P, k=1) - np.eye(P, k=-1)
        S[0, -1] = S[-1, 0] = -1
    return S * wn ** 2 / 2


@pytest.mark.parametrize("P", [1, 2, 5, 8, 10, 64])
def test_normal_modes_diagonalise_springs(oracle, P):
    wn = 0.7 * P
    U, w, _ = oracle.rp_normal_modes(P, wn, 0.1)
    assert np.allclose(U.T @ U, np.eye(P), atol=1e-13)
    assert np.isclose(abs(np.linalg.det(U)), 1.0)
    S = spring_matrix(P, wn)
    assert np.allclose(U.T @ S @ U, np.diag(w ** 2 / 2), atol=1e-10 * wn ** 2)
    assert np.allclose(w, 2 * wn * np.sin(np.arange(P) * np.pi / P))


def test_cayley_closed_form(oracle):
    P, wn, dt = 10, 3.0, 0.1
    _, w, Cm = oracle.rp_normal_modes(P, wn, dt)
    for k in range(P):
        A = np.array([[0.0, 1.0], [-w[k] ** 2, 0.0]])
        cay = np.linalg.inv(np.eye(2) - dt * A / 2) @ (np.eye(2) + dt * A / 2)
        assert np.allclose(Cm[k], cay, atol=1e-14)
        assert np.isclose(np.linalg.det(Cm[k]), 1.0)


def test_bcb_second_order_against_analytic_harmonic(oracle):
    """RPMD of a harmonic oscillator: each normal mode oscillates at sqrt(w_k^2 + w^2)."""
    P, m, w0, T = 10, 1.0, 1.0, 0.3
    model = oracle.make_model(oracle.MODEL_HARMONIC, 1, 1, [m, w0, 0.0])
    rng = np.random.default_rng(0)
    r0 = rng.normal(0, 0.3, size=(1, P)); v0 = rng.normal(0, 0.5, size=(1, P))
    U, w, _ = oracle.rp_normal_modes(P, P * T, 0.1)
    tend = 10.0
    Om = np.sqrt(w ** 2 + w0 ** 2)
    rn0, vn0 = U.T @ r0[0], U.T @ v0[0]
    exact = U @ (rn0 * np.cos(Om * tend) + vn0 / Om * np.sin(Om * tend))
    errs = []
    for dt in (0.0625, 0.03125, 0.015625):
        out = oracle.rp_run_ensemble(oracle.RP_CLASSICAL, model, m, r0, v0, dt=dt, nsteps=int(round(tend / dt)), temperature=T)
        errs.append(np.abs(out["r"][0] - exact).max())
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(1.6 < o < 2.4 for o in orders), (errs, orders)   # bcb.jl:44: order 2 atol 0.4


def test_rpmd_double_well_energy_conservation(oracle, key):
    P, m, beta = 16, 1836.0, 8.0
    model = oracle.make_model(oracle.MODEL_DW_GAUSSIAN, 1, 1, [1 / 1200, 1 / 40, 0.0, 0.0, 1.0])
    rng = np.random.default_rng(2)
    r0 = 3.87 + rng.normal(0, 0.1, size=(4, P)); v0 = rng.normal(0, np.sqrt(P / beta / m), size=(4, P))
    out = oracle.rp_run_ensemble(oracle.RP_CLASSICAL, model, m, r0, v0, dt=1.0, nsteps=2000, temperature=1 / beta)
    wn = P / beta

    def H(r, v):
        V = np.sum(r ** 4 / 1200 - r ** 2 / 40)
        spring = 0.5 * m * wn ** 2 * np.sum((r - np.roll(r, 1)) ** 2)
        return 0.5 * m * np.sum(v ** 2) + V + spring
    for i in range(4):
        assert np.isclose(H(out["r"][i], out["v"][i]), H(r0[i], v0[i]), rtol=2e-3)


@pytest.mark.parametrize("method", [1, 2])
def test_one_bead_ring_polymer_equals_plain_method(oracle, key, method):
    """BCBwithTsit5 with one bead is BABwithTsit5: RPSH / RP-Ehrenfest with P = 1 reproduce the
    FSSH / Ehrenfest oracle (centroid = the single bead, w_0 = 0)."""
    model = oracle.make_model(oracle.MODEL_DW_GAUSSIAN, 2, 1, [1 / 1200, 1 / 40, 0.01, 0.005, 1.0])
    ntraj = 64
    rng = np.random.default_rng(5)
    r0 = rng.normal(-0.5, 0.4, size=(ntraj, 1)); v0 = rng.normal(0.0, 0.01, size=(ntraj, 1))
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    st0 = np.zeros(ntraj, np.int32)
    rp = oracle.rp_run_ensemble(method, model, 1836.0, r0, v0, sig0, st0, dt=1.0, nsteps=500, temperature=0.125, key=key)
    pl = oracle.run_ensemble(0 if method == 1 else 1, model, 1836.0, r0, v0, sig0, state0=st0, dt=1.0, nsteps=500, key=key)
    assert np.allclose(rp["r"], pl["r"], rtol=1e-12, atol=1e-13)
    assert np.allclose(rp["v"], pl["v"], rtol=1e-12, atol=1e-15)
    assert np.array_equal(rp["state"], pl["state"])
    if method == 1:
        assert pl["nhops"].sum() > 0
