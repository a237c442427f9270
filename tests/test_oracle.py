"""Pins the FP64 CPU oracle (oracle/nqcd_oracle.cpp) before it is trusted as the
parity checker: known-answer vectors, the reference's own unit-test cases,
finite differences, closed-form solutions, invariants and the reference's
golden spin-boson curve (test/Dynamics/reference_data/gao_saller_jctc_2020_fig2b.csv).
"""
import os

import numpy as np
import pytest
from scipy.linalg import expm

HERE = os.path.dirname(os.path.abspath(__file__))
TULLY1 = [0.01, 1.6, 0.005, 1.0]
TULLY2 = [0.1, 0.28, 0.015, 0.06, 0.05]
TULLY3 = [6e-4, 0.1, 0.9]


# --- Philox: Random123 known-answer vectors (Salmon et al. SC'11) -----------------------
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_kat(oracle, ctr, key, expect):
    assert list(oracle.philox(ctr, key)) == expect


def test_uniform_and_normal_moments(oracle, key):
    u = np.array([oracle.uniform(i, 7, 1, key) for i in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    z = np.array([oracle.normal2(i, 0, 0, key) for i in range(10000)]).ravel()
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03


# --- Models: finite-difference derivatives (as test/Dynamics/utils.jl:39-65) ------------------
MODELS = [("tully1", 0, TULLY1), ("tully2", 1, TULLY2), ("tully3", 2, TULLY3), ("dw", 4, [1 / 1200, 1 / 40, 0.01, 0.005, 1.0])]


@pytest.mark.parametrize("name,kind,params", MODELS)
def test_model_derivative_finite_difference(oracle, name, kind, params):
    m = oracle.make_model(kind, 2, 1, params)
    for x in np.linspace(-4.3, 4.1, 23):
        h = 1e-6
        Vp = oracle.model_eval(m, [x + h])[0]
        Vm = oracle.model_eval(m, [x - h])[0]
        dV = oracle.model_eval(m, [x])[1][0]
        assert np.allclose(dV, (Vp - Vm) / (2 * h), atol=1e-8, rtol=1e-6), (name, x)


@pytest.mark.parametrize("n", [3, 4, 8])
def test_morse_n_model_finite_difference_and_closed_form(oracle, n):
    """n-state Morse model (Coker and Xiao form): dV by central differences, and the closed form at the
    Morse minimum x = r_i (V_ii = c_i) and far from every crossing (couplings vanish)."""
    import nstate_models as NS
    tab = NS.three_state() if n == 3 else NS.morse_table(n)
    m = oracle.make_model(oracle.MODEL_MORSE_N, n, 1, tab)
    for x in np.linspace(2.2, 9.0, 31):
        h = 1e-6
        Vp, Vm = oracle.model_eval(m, [x + h])[0], oracle.model_eval(m, [x - h])[0]
        V, dV = oracle.model_eval(m, [x])[:2]
        assert np.allclose(V, V.T) and np.allclose(dV[0], dV[0].T)
        assert np.allclose(dV[0], (Vp - Vm) / (2 * h), atol=1e-8, rtol=1e-6), x
    st = np.array(tab[:4 * n]).reshape(n, 4)
    for i in range(n):
        assert np.isclose(oracle.model_eval(m, [st[i, 2]])[0][i, i], st[i, 3], atol=1e-15)
    far = oracle.model_eval(m, [200.0])[0]  # every Gaussian coupling underflows to zero out here
    assert np.allclose(far - np.diag(np.diag(far)), 0.0, atol=1e-300)
    assert np.allclose(np.diag(far), st[:, 0] * (1 - np.exp(-st[:, 1] * (200.0 - st[:, 2]))) ** 2 + st[:, 3], rtol=1e-13)


# --- Spin-boson model (docs/src/NQCModels/systembathmodels.md:21-27) at eps != 0 ----------------
# V(R) = [[eps + c.R, Delta], [Delta, -eps - c.R]] + (1/2) sum_j w_j^2 R_j^2 (mass 1), so
# V(0) = [[eps, Delta], [Delta, -eps]] and the adiabatic energies are
# +-sqrt((eps + c.R)^2 + Delta^2) + (1/2) sum_j w_j^2 R_j^2.
SB_EPS, SB_DELTA = 1.0, 1.0


def _spin_boson(oracle, eps=SB_EPS, delta=SB_DELTA, nb=6, seed=11, c_scale=1.0):
    rng = np.random.default_rng(seed)
    omega = rng.uniform(0.2, 3.0, nb)
    c = c_scale * rng.uniform(0.1, 0.8, nb) * omega
    return oracle.make_model(3, 2, nb, [eps, delta], omega=omega, cbath=c), omega, c


def _spin_boson_fd_ok(model_eval, m, nb, rng):
    """Central differences of V and V0 against dV and dV0 at random bath coordinates."""
    for _ in range(8):
        R = rng.normal(0, 0.7, nb)
        V, dV, V0, dV0 = model_eval(m, R)
        for j in range(nb):
            h = 1e-6
            Rp, Rm = R.copy(), R.copy()
            Rp[j] += h; Rm[j] -= h
            Vp, _, V0p, _ = model_eval(m, Rp)
            Vm, _, V0m, _ = model_eval(m, Rm)
            if not np.allclose(dV[j], (Vp - Vm) / (2 * h), atol=1e-8, rtol=1e-6):
                return False
            if not np.isclose(dV0[j], (V0p - V0m) / (2 * h), atol=1e-8, rtol=1e-6):
                return False
    return True


def _spin_boson_closed_form_ok(model_eval, m, omega, c, eps, delta, rng):
    V, dV, V0, dV0 = model_eval(m, np.zeros(len(omega)))
    if not (np.allclose(V, [[eps, delta], [delta, -eps]], atol=1e-15) and V0 == 0.0 and np.all(dV0 == 0.0)):
        return False
    for _ in range(8):
        R = rng.normal(0, 0.7, len(omega))
        V, _, V0, _ = model_eval(m, R)
        bath = 0.5 * np.sum(omega ** 2 * R ** 2)
        gap = np.sqrt((eps + c @ R) ** 2 + delta ** 2)
        if not np.allclose(np.linalg.eigvalsh(V + V0 * np.eye(2)), [bath - gap, bath + gap], rtol=1e-13, atol=1e-14):
            return False
    return True


def test_spin_boson_derivative_finite_difference(oracle):
    """dV (state-dependent, per bath mode) and dV0 (the bath's harmonic force) by central
    differences at eps = 1 with nonzero couplings (as test/Dynamics/utils.jl:39-65)."""
    m, omega, c = _spin_boson(oracle)
    assert _spin_boson_fd_ok(oracle.model_eval, m, len(omega), np.random.default_rng(1))

    def planted_coupling_sign(mm, R):  # a sign error on c in the derivative only
        V, dV, V0, dV0 = oracle.model_eval(mm, R)
        return V, -dV, V0, dV0
    assert not _spin_boson_fd_ok(planted_coupling_sign, m, len(omega), np.random.default_rng(1))


def test_spin_boson_closed_form_values(oracle):
    """V(0) = [[eps, Delta], [Delta, -eps]] and eigenvalues +-sqrt((eps + c.R)^2 + Delta^2) + sum w^2 R^2 / 2;
    a model built with the sign of eps or of c flipped (what a sign error in either term of the
    oracle produces) fails the same check."""
    m, omega, c = _spin_boson(oracle)
    rng = lambda: np.random.default_rng(2)
    assert _spin_boson_closed_form_ok(oracle.model_eval, m, omega, c, SB_EPS, SB_DELTA, rng())
    m_eps, _, _ = _spin_boson(oracle, eps=-SB_EPS)
    assert not _spin_boson_closed_form_ok(oracle.model_eval, m_eps, omega, c, SB_EPS, SB_DELTA, rng())
    m_c, _, _ = _spin_boson(oracle, c_scale=-1.0)
    assert not _spin_boson_closed_form_ok(oracle.model_eval, m_c, omega, c, SB_EPS, SB_DELTA, rng())


def _rabi_run(oracle, key, eps, delta, dt, nsteps):
    """Ehrenfest on the spin-boson model with the bath uncoupled (c = 0), started in diabatic state 0:
    the diabatic population of state 0 at each step."""
    nb = 3
    omega = np.array([0.5, 1.0, 1.5])
    m = oracle.make_model(3, 2, nb, [eps, delta], omega=omega, cbath=np.zeros(nb))
    ntraj = 4
    r0, v0 = oracle.sample_gaussian(ntraj, nb, 0.0, 0.5, 0.0, 0.5, key)
    sig0 = oracle.diabatic_to_adiabatic(m, r0, 0)
    out = oracle.run_ensemble(1, m, 1.0, r0, v0, sig0, dt=dt, nsteps=nsteps, save_every=1, nsave=nsteps + 1, key=key)
    return out["pop_di"][:, :, 0]


def test_spin_boson_uncoupled_ehrenfest_rabi(oracle, key):
    """With c = 0 the two-level system is free: P0(t) = 1 - Delta^2/(eps^2 + Delta^2) sin^2(sqrt(eps^2 + Delta^2) tau).
    The electronic integrator's buffers start from a zero Hamiltonian at t = 0 (bab_electronics.jl:35-38,
    electronic_dynamics.jl:118-127), so the first step accumulates half a step of phase: tau = t - dt/2
    for t >= dt.  Pins Delta, |eps|, the adiabatic initial density and the diabatic population estimator."""
    dt, nsteps = 0.02, 500
    for eps, delta in ((SB_EPS, SB_DELTA), (0.5, 1.3)):
        P0 = _rabi_run(oracle, key, eps, delta, dt, nsteps)
        t = np.arange(nsteps + 1) * dt
        tau = np.where(t > 0, t - dt / 2, 0.0)
        W = np.sqrt(eps ** 2 + delta ** 2)
        exact = 1 - delta ** 2 / W ** 2 * np.sin(W * tau) ** 2
        assert np.abs(P0 - exact[None, :]).max() < 1e-8, np.abs(P0 - exact[None, :]).max()
        # the same law without the half-step of the zero start is far outside that bound
        assert np.abs(P0 - (1 - delta ** 2 / W ** 2 * np.sin(W * t) ** 2)[None, :]).max() > 1e-3


def test_diabatic_to_adiabatic_initial_density(oracle, key):
    """initialise_adiabatic_density_matrix (density_matrix_dynamics.jl:37-62,64-75): sigma_ad = U^T rho U for the
    diabatic pure state rho = e_d e_d^T, with U the eigenvectors at r0.  Checked against numpy's eigh brought to
    the sign rule pinned above (largest-magnitude component of each column positive): the diagonal U_da^2 (gauge
    free), the whole matrix, and the transform back, U sigma_ad U^T = rho."""
    m, omega, c = _spin_boson(oracle)
    ntraj, nb = 16, len(omega)
    r0, _ = oracle.sample_gaussian(ntraj, nb, 0.0, 0.8, 0.0, 0.1, key)
    rho = [np.diag([1.0, 0.0]), np.diag([0.0, 1.0])]
    for d in (0, 1):
        sig = oracle.diabatic_to_adiabatic(m, r0, d)
        for i in range(ntraj):
            _, Un = np.linalg.eigh(oracle.model_eval(m, r0[i])[0])
            Un = Un * np.sign(Un[np.argmax(np.abs(Un), axis=0), [0, 1]])[None, :]
            assert np.allclose(np.diag(sig[i]).real, Un[d, :] ** 2, atol=1e-14)
            assert np.allclose(sig[i], np.outer(Un[d], Un[d]), atol=1e-14)
            assert np.allclose(Un @ sig[i] @ Un.T, rho[d], atol=1e-14)
            assert np.allclose(sig[i].imag, 0) and np.isclose(np.trace(sig[i]).real, 1.0, atol=1e-15)
    # the other diabatic state's density (an index mistake) fails the back-transform (for two states U is a
    # symmetric reflection, so rows and columns coincide and a transposed U is no mistake here)
    _, Un = np.linalg.eigh(oracle.model_eval(m, r0[0])[0])
    assert not np.allclose(Un @ oracle.diabatic_to_adiabatic(m, r0[:1], 1)[0] @ Un.T, rho[0], atol=1e-6)


def test_model_known_values(oracle):
    t1 = oracle.make_model(0, 2, 1, TULLY1)
    V = oracle.model_eval(t1, [0.0])[0]
    assert np.allclose(V, [[0, 0.005], [0.005, 0]])          # crossing point: gap 2C
    V = oracle.model_eval(t1, [50.0])[0]
    assert np.allclose(V, [[0.01, 0], [0, -0.01]], atol=1e-12)
    t2 = oracle.make_model(1, 2, 1, TULLY2)
    V = oracle.model_eval(t2, [0.0])[0]
    assert np.allclose(V, [[0, 0.015], [0.015, -0.05]])
    t3 = oracle.make_model(2, 2, 1, TULLY3)
    assert np.isclose(oracle.model_eval(t3, [0.0])[0][0, 1], 0.1)
    assert np.isclose(oracle.model_eval(t3, [40.0])[0][0, 1], 0.2)
    assert np.isclose(oracle.model_eval(t3, [-40.0])[0][0, 1], 0.0, atol=1e-12)


# --- Eigen-decomposition and nonadiabatic coupling --------------------------------------
def test_eigen2_against_numpy_and_sign_rule(oracle):
    rng = np.random.default_rng(3)
    for _ in range(200):
        a, b, d = rng.normal(size=3)
        H = np.array([[a, b], [b, d]])
        E, U = oracle.eigen_sym(H)
        assert np.allclose(E, np.linalg.eigvalsh(H), atol=1e-14)
        assert np.allclose(U.T @ U, np.eye(2), atol=1e-14)
        assert np.allclose(U.T @ H @ U, np.diag(E), atol=1e-13)
        for j in range(2):  # largest-magnitude component positive
            assert U[np.argmax(np.abs(U[:, j])), j] > 0


def test_eigen_jacobi_against_numpy(oracle):
    rng = np.random.default_rng(4)
    for n in (3, 5, 8):
        A = rng.normal(size=(n, n))
        H = A + A.T
        E, U = oracle.eigen_sym(H)
        assert np.allclose(E, np.linalg.eigvalsh(H), atol=1e-12)
        assert np.allclose(U.T @ H @ U, np.diag(E), atol=1e-11)


def test_eigen_tracking_flips_to_previous(oracle):
    H = np.array([[0.3, 0.1], [0.1, -0.2]])
    E, U = oracle.eigen_sym(H)
    E2, U2 = oracle.eigen_sym(H, -U)
    assert np.allclose(U2, -U)


@pytest.mark.parametrize("name,kind,params", MODELS[:3])
def test_nac_antisymmetric_and_matches_eigenvector_derivative(oracle, name, kind, params):
    """d = -d^T (test/Dynamics/fssh.jl:33) and d_ij = <i|d/dx j> by finite differences."""
    m = oracle.make_model(kind, 2, 1, params)
    for x in (-1.3, -0.2, 0.4, 1.7):
        V, dV, _, _ = oracle.model_eval(m, [x])
        E, U = oracle.eigen_sym(V)
        G, d = oracle.adiabatic_quantities(E, U, dV)
        assert np.allclose(d[0], -d[0].T)
        h = 1e-6
        _, Up = oracle.eigen_sym(oracle.model_eval(m, [x + h])[0], U)
        _, Um = oracle.eigen_sym(oracle.model_eval(m, [x - h])[0], U)
        dU = (Up - Um) / (2 * h)
        assert np.allclose(d[0], U.T @ dU, atol=1e-6 * max(1, abs(d[0]).max())), (name, x)
        # Hellmann-Feynman: diagonal of U^T dV U equals dE/dx
        Ep = oracle.eigen_sym(oracle.model_eval(m, [x + h])[0])[0]
        Em = oracle.eigen_sym(oracle.model_eval(m, [x - h])[0])[0]
        assert np.allclose(np.diag(G[0]), (Ep - Em) / (2 * h), atol=1e-8)


# --- Density-matrix propagation (electronic_dynamics.jl, Tsit5) ---------------------------
def _herm(rng, n):
    c = rng.normal(size=n) + 1j * rng.normal(size=n)
    c /= np.linalg.norm(c)
    return np.outer(c, c.conj())


def test_density_constant_hamiltonian_matches_exponential(oracle):
    rng = np.random.default_rng(5)
    n = 2
    E = np.array([-0.01, 0.02]); w = 0.003
    W = np.array([[0, w], [-w, 0]])
    V = np.diag(E) - 1j * W
    sig0 = _herm(rng, n)
    dt = 1.0
    sig = oracle.propagate_density(sig0, E, W, 0.0, E, W, dt, 0.0, dt)
    U = expm(-1j * V * dt)
    exact = U @ sig0 @ U.conj().T
    assert np.allclose(sig, exact, atol=1e-13)
    assert np.isclose(np.trace(sig).real, 1.0, atol=1e-14)
    assert np.allclose(sig, sig.conj().T, atol=1e-15)


def test_density_fifth_order_convergence(oracle):
    rng = np.random.default_rng(6)
    E0 = np.array([0.0, 0.6]); E1 = np.array([-0.2, 0.9])
    W0 = np.array([[0, 0.3], [-0.3, 0]]); W1 = np.array([[0, -0.5], [0.5, 0]])
    sig0 = _herm(rng, 2)
    ref = oracle.propagate_density(sig0, E0, W0, 0.0, E1, W1, 2.0, 0.0, 2.0, nsub=400)
    errs = [np.abs(oracle.propagate_density(sig0, E0, W0, 0.0, E1, W1, 2.0, 0.0, 2.0, nsub=k) - ref).max() for k in (5, 10, 20)]
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    # global order >= 5 (linear problems show ~6); a wrong coefficient drops it to <= 4
    assert all(4.5 < o < 6.5 for o in orders), (errs, orders)


def test_density_trace_preserved_many_states(oracle):
    """DensityMatrixODEProblem norm preservation (test/Dynamics/electronic_dynamics.jl:35-47)."""
    rng = np.random.default_rng(7)
    n = 10
    A = rng.normal(size=(n, n)); A = A @ A.T
    A = np.triu(A, 1); W = A - A.T  # antisymmetric couplings
    E = rng.random(n)
    sig0 = np.zeros((n, n), complex); sig0[0, 0] = 1
    sig = oracle.propagate_density(sig0, E, W, 0.0, E, W, 1.0, 0.0, 0.5, nsub=50)
    assert np.isclose(np.trace(sig).real, 1.0, atol=1e-12)


def test_density_linear_interpolation_from_zero_buffer(oracle):
    """First step interpolates from a zero Hamiltonian at t=0 (electronic_dynamics.jl:118-127):
    for diagonal V the coherence picks up exp(-i int_0^dt dE(t) dt) = exp(-i dE dt / 2)."""
    dE = 0.8
    E1 = np.array([0.5 * dE, -0.5 * dE])
    Z = np.zeros((2, 2))
    sig0 = np.array([[0.5, 0.5], [0.5, 0.5]], complex)
    sig = oracle.propagate_density(sig0, np.zeros(2), Z, 0.0, E1, Z, 1.0, 0.0, 1.0, nsub=5)
    assert np.isclose(sig[0, 1], 0.5 * np.exp(-1j * dE * 0.5), atol=1e-9)


# --- Hopping: probability, selection, rescaling --------------------------------------------
def test_select_new_state_known_answers(oracle):
    """test/Dynamics/fssh.jl:53-63 (states 0-based here)."""
    for r in (0.0, 0.5, 0.999999):
        assert oracle.fssh_select([0.0, 1.0], 0, r) == 1
        assert oracle.fssh_select([0.0, 1.0], 1, r) == 1
        assert oracle.fssh_select([1.0, 0.0], 1, r) == 0
    assert oracle.fssh_select([0.0, 0.3], 0, 0.3) == 0   # strict prob > r (fssh.jl:115)


def test_probability_equals_population_flux(oracle):
    """Fewest switches: g_{s->m} dt-limit equals -d(sigma_ss)/sigma_ss over the step, fixing the
    sign and index conventions of d[s,m] and sigma[m,s] (fssh.jl:96-108)."""
    rng = np.random.default_rng(8)
    for trial in range(20):
        sig0 = _herm(rng, 2)
        E = np.array([-0.01, 0.01])
        v = np.array([rng.normal() * 0.01])
        dval = rng.normal()
        d = np.array([[[0.0, dval], [-dval, 0.0]]])
        W = d[0] * v[0]
        h = 1e-5
        for s in (0, 1):
            prob = oracle.fssh_probability(v, sig0, s, d, h)
            sig = oracle.propagate_density(sig0, E, W, 0.0, E, W, h, 0.0, h, nsub=1)
            flux = -(sig[s, s].real - sig0[s, s].real) / sig0[s, s].real
            m = 1 - s
            pm = prob[m] - (prob[m - 1] if m > 0 else 0.0)
            assert np.isclose(pm, max(flux, 0.0), rtol=1e-3, atol=1e-12), (trial, s, pm, flux)


def test_probability_clamped(oracle):
    sig = np.array([[0.5, 0.5], [0.5, 0.5]], complex)
    d = np.array([[[0.0, 1e6], [-1e6, 0.0]]])
    assert np.allclose(oracle.fssh_probability([1.0], sig, 0, d, 1.0), [0.0, 1.0])
    assert np.allclose(oracle.fssh_probability([-1.0], sig, 0, d, 1.0), [0.0, 0.0])


def test_rescale_accept_reject_and_energy(oracle):
    """test/Dynamics/fssh.jl:65-74 (frustrated vs accepted) and :87-120 (dKE = -dE)."""
    mass = np.array([2000.0, 1500.0])
    E = np.array([-0.01, 0.02])
    d = np.zeros((2, 2, 2)); d[0, 1, 0], d[1, 1, 0] = 0.7, -0.4; d[:, 0, 1] = -d[:, 1, 0]
    ok, v = oracle.rescale_velocity([0.0, 0.0], mass, E, d, 1, 0)
    assert not ok and np.allclose(v, 0)
    v0 = np.array([0.02, -0.01])
    ok, v = oracle.rescale_velocity(v0, mass, E, d, 1, 0)
    assert ok
    dKE = 0.5 * np.sum(mass * v ** 2) - 0.5 * np.sum(mass * v0 ** 2)
    assert np.isclose(dKE, -(E[1] - E[0]), rtol=1e-12)
    # the velocity change is along d / m
    dv = v - v0
    assert np.isclose(dv[0] / dv[1], (d[0, 1, 0] / mass[0]) / (d[1, 1, 0] / mass[1]))
    # vinversion reflects the component along d and keeps |v|
    vs = np.array([1e-4, 2e-4])
    ok, v = oracle.rescale_velocity(vs, mass, E, d, 1, 0, rescaling=1)
    dn = d[:, 1, 0] / np.linalg.norm(d[:, 1, 0])
    assert not ok and np.isclose(np.linalg.norm(v), np.linalg.norm(vs)) and np.isclose(v @ dn, -(vs @ dn))
    ok, v = oracle.rescale_velocity(vs, mass, E, d, 1, 0, rescaling=2)
    assert ok and np.allclose(v, vs)


# --- Whole trajectories -------------------------------------------------------------------
def _energy_fssh(oracle, m, mass, r, v, s):
    E, _ = oracle.eigen_sym(oracle.model_eval(m, r)[0])
    return 0.5 * mass * np.sum(v ** 2) + E[s]


def test_fssh_energy_conservation_tully2(oracle, key):
    """test/Dynamics/fssh.jl:123-131: Tully 2, x=-10, v=100/2000, dt=0.1, 500 a.u., rtol 1e-2."""
    m = oracle.make_model(1, 2, 1, TULLY2)
    r0 = np.array([[-10.0]]); v0 = np.array([[100 / 2000]])
    sig0 = np.zeros((1, 2, 2), complex); sig0[0, 0, 0] = 1
    out = oracle.run_ensemble(0, m, 2000.0, r0, v0, sig0, state0=[0], dt=0.1, nsteps=5000, key=key)
    e0 = _energy_fssh(oracle, m, 2000.0, r0[0], v0[0], 0)
    e1 = _energy_fssh(oracle, m, 2000.0, out["r"][0], out["v"][0], out["state"][0])
    assert np.isclose(e0, e1, rtol=1e-2)


def test_fssh_hops_conserve_energy_and_trace(oracle, key):
    m = oracle.make_model(0, 2, 1, TULLY1)
    ntraj = 400
    r0, v0 = oracle.sample_gaussian(ntraj, 1, -3.0, 0.5, 15 / 2000, 0.0, key)
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    out = oracle.run_ensemble(0, m, 2000.0, r0, v0, sig0, state0=np.zeros(ntraj, np.int32), dt=1.0, nsteps=800, key=key)
    assert out["nhops"].sum() > 50
    rel = np.zeros(ntraj)
    for i in range(ntraj):
        e0 = _energy_fssh(oracle, m, 2000.0, r0[i], v0[i], 0)
        e1 = _energy_fssh(oracle, m, 2000.0, out["r"][i], out["v"][i], out["state"][i])
        rel[i] = abs(e1 - e0) / abs(e0)
    # without hops velocity Verlet conserves tightly; the half-kick after a hop reuses the
    # pre-hop force (bab_electronics.jl:72,79), an O(dt v dF) error bounded by the
    # reference at rtol 1e-2 per run (fssh.jl:131)
    assert np.all(rel[out["nhops"] == 0] < 1e-5)
    assert np.all(rel < 1e-2)
    tr = out["sig_re"][:, 0, 0] + out["sig_re"][:, 1, 1]
    assert np.allclose(tr, 1.0, atol=1e-10)


def test_ehrenfest_energy_conservation(oracle):
    """test/Dynamics/ehrenfest.jl:45-57: Ehrenfest on Tully 2 conserves KE + sum_i sigma_ii E_i."""
    m = oracle.make_model(1, 2, 1, TULLY2)
    r0 = np.array([[-10.0]]); v0 = np.array([[100 / 2000]])
    sig0 = np.zeros((1, 2, 2), complex); sig0[0, 0, 0] = 1
    for nsteps in (2000, 20000):
        out = oracle.run_ensemble(1, m, 2000.0, r0, v0, sig0, dt=0.01, nsteps=nsteps)
        E, _ = oracle.eigen_sym(oracle.model_eval(m, out["r"][0])[0])
        e1 = 0.5 * 2000 * out["v"][0, 0] ** 2 + out["sig_re"][0, 0, 0] * E[0] + out["sig_re"][0, 1, 1] * E[1]
        E0, _ = oracle.eigen_sym(oracle.model_eval(m, r0[0])[0])
        e0 = 0.5 * 2000 * v0[0, 0] ** 2 + E0[0]
        assert abs(e1 - e0) < 1e-5 * e0  # reference: var(E) ~ 0 atol 1e-6 (ehrenfest.jl:56)


def test_tully1_fssh_transmission_symmetric_limit(oracle, key):
    """High-momentum Tully 1 (k=30): everything transmits; the fractions on each surface are
    bounded by the Landau-Zener two-state estimate of hopping at the single crossing."""
    m = oracle.make_model(0, 2, 1, TULLY1)
    ntraj = 500
    r0, v0 = oracle.sample_gaussian(ntraj, 1, -5.0, 30 / (20 * np.sqrt(2)), 30 / 2000, 0.0, key)
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    out = oracle.run_ensemble(0, m, 2000.0, r0, v0, sig0, state0=np.zeros(ntraj, np.int32), dt=1.0, nsteps=1500, key=key)
    assert np.all(out["r"][:, 0] > 0)
    v = 30 / 2000
    alpha = 2 * np.pi * 0.005 ** 2 / (v * 2 * 0.01 * 1.6)  # LZ: 2 pi V12^2 / (hbar v |d(V11-V22)/dx|)
    p_lz_stay_diabatic = np.exp(-alpha)  # diabatic passage = adiabatic hop
    p_upper = np.mean(out["state"] == 1)
    assert abs(p_upper - p_lz_stay_diabatic) < 0.12


# --- Golden curve: spin-boson Ehrenfest (test/Dynamics/ehrenfest.jl:110-141) ------------------
def test_spin_boson_ehrenfest_gao_saller(oracle, key):
    N = 100; beta = 5.0; wc = 2.5; alpha = 0.09
    j = np.arange(1, N + 1)
    omega = -wc * np.log(1 - j / (N + 1))
    c = np.sqrt(alpha * wc / (N + 1)) * omega
    m = oracle.make_model(3, 2, N, [0.0, 1.0], omega=omega, cbath=c)
    ntraj = 500
    t = np.tanh(beta * omega / 2)
    sd_r = np.sqrt(1 / (2 * omega * t)); sd_v = np.sqrt(omega / (2 * t))
    r0, v0 = oracle.sample_gaussian(ntraj, N, 0.0, sd_r[None, :], 0.0, sd_v[None, :], key)
    sig0 = oracle.diabatic_to_adiabatic(m, r0, 0)
    out = oracle.run_ensemble(1, m, 1.0, r0, v0, sig0, dt=0.1, nsteps=200, save_every=1, nsave=201, key=key)
    P = out["pop_di"]  # (ntraj, nsave, 2)
    corr = np.mean(P[:, 0:1, 0] * P[:, :, 0] - P[:, 0:1, 0] * P[:, :, 1], axis=0)
    data = np.loadtxt(os.path.join(HERE, "golden", "gao_saller_jctc_2020_fig2b.csv"), delimiter=",")
    order = np.argsort(data[:, 0])
    tg, yg = data[order, 0], data[order, 1]
    times = np.arange(201) * 0.1
    # linear interpolation with linear extrapolation (Interpolations.Line())
    ref = np.interp(times, tg, yg)
    lo = times < tg[0]; hi = times > tg[-1]
    ref[lo] = yg[0] + (times[lo] - tg[0]) * (yg[1] - yg[0]) / (tg[1] - tg[0])
    ref[hi] = yg[-1] + (times[hi] - tg[-1]) * (yg[-1] - yg[-2]) / (tg[-1] - tg[-2])
    assert np.all(np.abs(corr - ref) < 0.2), np.abs(corr - ref).max()


@pytest.mark.parametrize("method", [0, 1])
def test_saved_energies(oracle, key, method):
    """OutputKineticEnergy / OutputTotalEnergy (DynamicsOutputs.jl:90-98): the first and last save slots
    equal m v^2 / 2 and classical_potential_energy computed independently from the initial and final
    positions (numpy eigenvalues; FSSH: the active state's, fssh.jl:150-154; Ehrenfest: sum_i
    sigma_ii E_i, ehrenfest.jl:85-95), and the total is conserved over the slots within the reference's
    rtol of fssh.jl:131 (dt = 1 through the crossing; a hop's half-kick reuses the old force)."""
    m = oracle.make_model(0, 2, 1, TULLY1)
    ntraj = 64
    r0, v0 = oracle.sample_gaussian(ntraj, 1, -4.0, 0.5, 15 / 2000, 0.0, key)
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    out = oracle.run_ensemble(method, m, 2000.0, r0, v0, sig0, state0=np.zeros(ntraj, np.int32), dt=1.0, nsteps=600,
                              key=key, save_every=50, nsave=13, energies=True)
    for i in range(ntraj):
        E0 = np.linalg.eigvalsh(oracle.model_eval(m, r0[i])[0])
        E1 = np.linalg.eigvalsh(oracle.model_eval(m, out["r"][i])[0])
        assert np.isclose(out["ekin"][i, 0], 0.5 * 2000 * v0[i, 0] ** 2, rtol=1e-14)
        assert np.isclose(out["epot"][i, 0], E0[0], rtol=1e-12, atol=1e-15)
        assert np.isclose(out["ekin"][i, -1], 0.5 * 2000 * out["v"][i, 0] ** 2, rtol=1e-14)
        p1 = E1[out["state"][i]] if method == 0 else out["sig_re"][i, 0, 0] * E1[0] + out["sig_re"][i, 1, 1] * E1[1]
        assert np.isclose(out["epot"][i, -1], p1, rtol=1e-12, atol=1e-15)
    tot = out["ekin"] + out["epot"]
    rel = np.abs(tot - tot[:, :1]) / np.abs(tot[:, :1])
    assert rel.max() < 1e-2
    if method == 0:
        assert out["nhops"].sum() > 0


def test_termination_is_evaluated_after_each_step(oracle, key):
    """TerminatingCallback is a DiscreteCallback (DynamicsUtils/callbacks.jl:15-29): its condition is
    checked after every step and not at t0, so a trajectory that starts past |x| > 10 takes exactly
    one step, and one that reaches the boundary stops at the first step after crossing it."""
    m = oracle.make_model(0, 2, 1, TULLY1)
    r0 = np.array([[11.0], [-12.0], [9.985]])
    v0 = np.array([[0.01], [-0.01], [0.01]])
    sig0 = np.zeros((3, 2, 2), complex); sig0[:, 0, 0] = 1
    out = oracle.run_ensemble(0, m, 2000.0, r0, v0, sig0, state0=np.zeros(3, np.int32), dt=1.0, nsteps=100, key=key,
                              term_x=10.0)
    assert list(out["steps"][:2]) == [1, 1]
    assert np.isclose(out["r"][0, 0], 11.0 + 0.01, atol=1e-6) and np.isclose(out["r"][1, 0], -12.0 - 0.01, atol=1e-6)
    assert out["steps"][2] == 2 and out["r"][2, 0] > 10.0  # 9.985 -> 9.995 -> 10.005
