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


# --- path-integral Monte Carlo (MetropolisHastings.jl:131-148,170-258) ---------------------------
def _hartley(P):
    j = np.arange(P)
    a = 2 * np.pi * ((j[:, None] * j[None, :]) % P) / P
    return (np.cos(a) + np.sin(a)) / np.sqrt(P)


def _pimc_mode_variances(P, mass, w, T, width):
    """Stationary normal-mode variances of the chain without the second test (a Metropolis chain): the
    target exp(-beta_P sum_b V(x_b)) times the resampling laws N(0, sigma_k) of modes 1..P-1 (the proposals
    are independence draws accepted on the potential alone), for V = m w^2 x^2 / 2 Gaussian in every mode."""
    wn = P * T
    wk = 2 * wn * np.sin(np.arange(P) * np.pi / P)
    sig = np.zeros(P)
    if P > 1:
        sig[1:] = np.sqrt(wn / mass) / wk[1:] if width else np.sqrt(wn / mass) / (0.5 * wk[1:] ** 2 * mass)
    prec = mass * w * w / wn + np.r_[0.0, 1.0 / sig[1:] ** 2]
    return 1.0 / prec


@pytest.mark.parametrize("P,width", [(8, 1), (8, 0), (4, 1), (16, 1)])
def test_pimc_stationary_law_harmonic(oracle, P, width):
    """Independent chains on a harmonic well: the centroid variance 1/(beta_P m w^2) pins the bead-summed
    potential and the temperature w_n = P T of the acceptance; the other modes pin sigma_k and the mode
    frequencies.  With width = 1 (the free ring polymer's sigma_k) this is the discretised path-integral
    distribution, variance 1/(beta_P m (w^2 + w_k^2)) per mode.  5 sigma of the variance estimate."""
    mass, w, T = 1.3, 0.9, 1.0 / 8
    m = oracle.make_model(oracle.MODEL_HARMONIC, 1, 1, [mass, w, 0.0])
    ntraj = 20000
    x, v, acc = oracle.pimc(m, np.zeros((ntraj, P)), mass, T, delta=1.5, passes=300, width=width, key=11, reassess=0)
    y = x @ _hartley(P).T
    exp_var = _pimc_mode_variances(P, mass, w, T, width)
    z = (y.var(axis=0) - exp_var) / (exp_var * np.sqrt(2.0 / ntraj))
    assert np.all(np.abs(z) < 5), z
    assert np.abs(y.mean(axis=0)).max() < 5 * np.sqrt(exp_var.max() / ntraj)
    if width:  # the free-ring-polymer width gives the path-integral law of the harmonic sampler
        wk = 2 * P * T * np.sin(np.arange(P) * np.pi / P)
        assert np.allclose(exp_var, 1 / (mass * (w ** 2 + wk ** 2) / (P * T)))
    # velocities: N(0, sqrt(w_n / m)) per bead
    assert abs(v.var() / (P * T / mass) - 1) < 5 * np.sqrt(2.0 / v.size)
    assert 0 < acc.mean() < 300 * (2 + (1 if P > 1 else 0))


def _p1_stationary_var(mass, w, T, delta, reassess, nx=1601, xmax=6.0):
    """Exact stationary variance of the one-bead chain by its transition kernel on a grid: a uniform
    proposal of width delta accepted with probability a = min(1, exp(-dV/T)), or 1 - (1 - a)^2 when the
    proposal is tested twice, as run_main_loop! does (MetropolisHastings.jl:194)."""
    x = np.linspace(-xmax, xmax, nx)
    dx = x[1] - x[0]
    V = 0.5 * mass * w * w * x * x
    a = np.minimum(1.0, np.exp(-(V[None, :] - V[:, None]) / T))
    g = 1 - (1 - a) ** 2 if reassess else a
    K = np.where(np.abs(x[None, :] - x[:, None]) < delta / 2, g * dx / delta, 0.0)
    K[np.arange(nx), np.arange(nx)] += 1 - K.sum(axis=1)
    p = np.exp(-V / T)
    p /= p.sum()
    for _ in range(4000):
        p = p @ K
    return float((p * x * x).sum())


@pytest.mark.parametrize("delta", [1.0, 3.0])
def test_pimc_second_test_of_last_proposal(oracle, delta):
    """The reference tests the last proposal of each step twice (MetropolisHastings.jl:194), so a rejected
    proposal gets a second chance and the chain is no longer a Metropolis chain: on one bead its law is
    the stationary law of the kernel with acceptance 1 - (1 - a)^2, computed here on a grid (5 sigma), and
    differs from the Boltzmann law T/(m w^2) that the chain without it reaches."""
    mass, w, T = 1.0, 1.0, 0.5
    m = oracle.make_model(oracle.MODEL_HARMONIC, 1, 1, [mass, w, 0.0])
    ntraj = 40000
    for reassess in (1, 0):
        x, _, _ = oracle.pimc(m, np.zeros((ntraj, 1)), mass, T, delta=delta, passes=400, width=1, key=7,
                              reassess=reassess)
        ref = _p1_stationary_var(mass, w, T, delta, reassess)
        assert abs(x.var() - ref) < 5 * ref * np.sqrt(2.0 / ntraj), (reassess, x.var(), ref)
        if reassess:
            assert ref > 1.3 * T / (mass * w * w)
        else:
            assert abs(ref - T / (mass * w * w)) < 1e-6
