"""Pins the IESH / EhrenfestNA part of the oracle: Hermitian eigen-solver (zheevr's role),
LU determinant (test/Core/FastDeterminant.jl:9-13), the Newns-Anderson matrices,
unitarity of the wavefunction propagation, hop energy conservation (test/Dynamics/iesh.jl:122-151)
and EhrenfestNA energy conservation."""
import numpy as np
import pytest


def test_hermitian_eigen_against_numpy(oracle):
    rng = np.random.default_rng(1)
    for n in (2, 5, 17, 41):
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        H = A + A.conj().T
        E, U = oracle.eigen_hermitian(H)
        assert np.allclose(E, np.linalg.eigvalsh(H), atol=1e-11)
        assert np.allclose(U.conj().T @ U, np.eye(n), atol=1e-12)
        assert np.allclose(U.conj().T @ H @ U, np.diag(E), atol=1e-10)


def test_lu_determinant_against_numpy(oracle):
    rng = np.random.default_rng(2)
    for n in (1, 3, 10, 25):
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        assert np.isclose(oracle.det(A), np.linalg.det(A), rtol=1e-11)


def test_newns_anderson_matrices(oracle):
    md = oracle.na_model(M=20)
    for x in (-5.0, 3.0, 21.0):
        E, U, dh = oracle.na_eigen(md, x)
        N = 21
        W = 6 * 6.4e-3 / 2
        eps = -W + np.arange(20) * 2 * W / 19
        Vk = np.sqrt(6.4e-3 / (2 * np.pi)) * np.sqrt(2 * W / 19)
        mw2 = 2000 * 2e-4 ** 2
        h = 0.5 * mw2 * (x - 20.6097) ** 2 - 3.8e-3 - 0.5 * mw2 * x ** 2
        H = np.diag(np.concatenate([[h], eps])); H[0, 1:] = H[1:, 0] = Vk
        assert np.allclose(E, np.linalg.eigvalsh(H), atol=1e-14)
        assert np.allclose(U.T @ H @ U, np.diag(E), atol=1e-13)
        assert np.isclose(dh, -mw2 * 20.6097)


def _na_energy(md, mass, x, v, occ=None, psi=None):
    oracle = __import__("oracle")
    E, _, _ = oracle.na_eigen(md, x)
    e = 0.5 * mass * v ** 2 + 0.5 * md.mw2 * x ** 2
    if occ is not None:
        return e + E[occ].sum()
    return e + np.sum(E[:, None] * np.abs(psi) ** 2)


def test_iesh_unitarity_and_hop_energy(oracle, key):
    md = oracle.na_model(M=30)
    N, Ne, mass = 31, 15, 2000.0
    ntraj = 24
    rng = np.random.default_rng(3)
    # the impurity level crosses the Fermi level near x = 8
    r0 = rng.normal(8.0, 2.0, size=ntraj); v0 = np.abs(rng.normal(0.0, 1e-2, size=ntraj))
    occ0 = np.tile(np.arange(Ne), (ntraj, 1))
    occ0[:, -1] = Ne  # one electron above the Fermi level
    out = oracle.iesh_run_ensemble(0, md, Ne, mass, r0, v0, occ0, dt=2.0, nsteps=400, key=key)
    assert out["nhops"].sum() > 5
    # the early-exit estimate (iesh.jl:241-244) bounds the total probability, so it changes nothing
    out2 = oracle.iesh_run_ensemble(0, md, Ne, mass, r0, v0, occ0, dt=2.0, nsteps=400, key=key, estimate=False)
    assert np.array_equal(out["occ"], out2["occ"]) and np.array_equal(out["r"], out2["r"])
    for i in range(ntraj):
        psi = out["psi"][i]
        assert np.allclose(psi.conj().T @ psi, np.eye(Ne), atol=1e-9)      # propagation is unitary
        e0 = _na_energy(md, mass, r0[i], v0[i], occ0[i])
        e1 = _na_energy(md, mass, out["r"][i], out["v"][i], out["occ"][i])
        assert abs(e1 - e0) < 1e-2 * abs(e0) + 2e-4                          # rtol 1e-2 (iesh.jl test)


def test_ehrenfest_na_energy_conservation(oracle, key):
    md = oracle.na_model(M=20)
    Ne, mass = 10, 2000.0
    r0 = np.array([21.0, 15.0]); v0 = np.array([0.0, 1e-3])
    occ0 = np.tile(np.arange(Ne), (2, 1))
    out = oracle.iesh_run_ensemble(1, md, Ne, mass, r0, v0, occ0, dt=1.0, nsteps=2000, key=key)
    for i in range(2):
        psi0 = np.zeros((21, Ne)); psi0[np.arange(Ne), np.arange(Ne)] = 1
        e0 = _na_energy(md, mass, r0[i], v0[i], psi=psi0)
        e1 = _na_energy(md, mass, out["r"][i], out["v"][i], psi=out["psi"][i])
        assert abs(e1 - e0) < 1e-6


def test_hop_probability_determinant_lemma(oracle):
    """det(S with row n replaced by psi_m) = det(S) (psi_m . S^-1[:, n]) -- the identity the device
    kernel uses in place of one LU per candidate (iesh.jl:256-266,301-306)."""
    rng = np.random.default_rng(4)
    N, Ne = 12, 5
    A = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    Q, _ = np.linalg.qr(A)
    psi = Q[:, :Ne]
    occ = np.array([0, 2, 3, 7, 9])
    S = psi[occ, :]
    Sinv = np.linalg.inv(S)
    d0 = oracle.det(S)
    for n in range(Ne):
        for m in (1, 4, 11):
            Sn = S.copy(); Sn[n, :] = psi[m, :]
            assert np.isclose(oracle.det(Sn), d0 * (psi[m, :] @ Sinv[:, n]), rtol=1e-9)


def test_edc_keeps_each_electron_normalised(oracle, key):
    md = oracle.na_model(M=20)
    Ne, mass, ntraj = 10, 2000.0, 8
    rng = np.random.default_rng(5)
    r0 = rng.normal(8.0, 1.0, size=ntraj); v0 = np.abs(rng.normal(0.0, 1e-2, size=ntraj))
    occ0 = np.tile(np.arange(Ne), (ntraj, 1))
    out = oracle.iesh_run_ensemble(0, md, Ne, mass, r0, v0, occ0, dt=2.0, nsteps=200, key=key, edc=True)
    norms = np.sum(np.abs(out["psi"]) ** 2, axis=1)
    assert np.allclose(norms, 1.0, atol=1e-12)


@pytest.mark.parametrize("h_over_gamma", [-1.0, 0.0, 0.5, 2.0])
def test_newns_anderson_impurity_occupation_wide_band(oracle, h_over_gamma):
    """Thermal impurity occupation of the discretised Newns-Anderson model against the wide-band
    Lorentzian n_d = int rho(w) f(w) dw, rho = (Gamma/2pi) / ((w - h)^2 + (Gamma/2)^2), Fermi level 0
    (at T = 0 this is 1/2 - atan(2h/Gamma)/pi).  Pins the bath coupling V_k = sqrt(Gamma Delta_eps / 2pi)."""
    from scipy.integrate import quad
    Gamma = 6.4e-3
    md = oracle.na_model(Gamma=Gamma, M=400, W=20 * Gamma)
    kT = 2 * (md.b - md.a) / (md.M - 1)          # a few level spacings smooths the discretisation
    h = h_over_gamma * Gamma
    x = (0.5 * md.mw2 * md.g ** 2 + md.dG - h) / (md.mw2 * md.g)
    E, U, _ = oracle.na_eigen(md, x)
    f = lambda w: 0.5 * (1 - np.tanh(w / (2 * kT)))
    nd = np.sum(U[0] ** 2 * f(E))
    rho = lambda w: (Gamma / (2 * np.pi)) / ((w - h) ** 2 + (Gamma / 2) ** 2)
    L = 400 * Gamma
    expected = quad(lambda w: rho(w) * f(w), -L, L, points=[0.0, h], limit=400)[0] + quad(rho, -np.inf, -L)[0]
    if h_over_gamma == 0.0:
        assert abs(expected - 0.5) < 1e-9
    assert abs(nd - expected) < 0.01          # finite band (W = 20 Gamma): |error| <= 0.006; Gamma off by 2x: 0.08


def _na_nac(oracle, md, x):
    E, U, dh = oracle.na_eigen(md, x)
    U = U * np.sign(U[0])  # impurity-positive gauge
    G = dh * np.outer(U[0], U[0])
    with np.errstate(divide="ignore"):
        d = G / (E[None, :] - E[:, None])
    np.fill_diagonal(d, 0.0)
    return E, U, d


def test_iesh_probability_is_the_determinant_population_flux(oracle):
    """Fewest switches for Slater determinants: the unclamped probabilities summed over (n, m) equal
    -dt d|det S|^2/dt / |det S|^2 under i dpsi/dt = (diag(E) - i v d) psi.  Pins the prefactor, the
    sign of v_dot_d and the index order d[m, occ_n] (iesh.jl:256-266,285-298)."""
    from scipy.linalg import expm
    md = oracle.na_model(M=12)
    E, U, d = _na_nac(oracle, md, 7.5)
    N, Ne, v, dt = 13, 5, 3e-3, 0.7
    occ = np.array([0, 2, 3, 5, 6])
    rng = np.random.default_rng(21)
    X = np.eye(N)[:, occ] + 0.4 * (rng.normal(size=(N, Ne)) + 1j * rng.normal(size=(N, Ne)))
    psi, _ = np.linalg.qr(X)
    H = np.diag(E) - 1j * v * d
    assert np.allclose(H, H.conj().T)
    A = lambda p: abs(np.linalg.det(p[occ, :])) ** 2
    h = 1e-2
    dA = (A(expm(-1j * H * h) @ psi) - A(expm(1j * H * h) @ psi)) / (2 * h)
    prob = oracle.iesh_probabilities(psi, occ, d, v, dt, clamp=False)
    assert np.all(prob[occ, :] == 0.0)
    assert abs(prob.sum() - (-dt * dA / A(psi))) < 1e-7 * max(1.0, abs(prob).sum())
    assert abs(prob).sum() > 1e-4


def test_ehrenfest_na_against_exact_equations_of_motion(oracle):
    """EhrenfestNA with VerletwithElectronics at dt = 10 against a tight Runge-Kutta solution of
    dx/dt = v, m dv/dt = -dU0/dx - sum_e psi_e^H (dh u u^T) psi_e, i dpsi/dt = (diag(E) - i v d) psi
    (the reference's own check against Vern9 at rtol 1e-3, test/Dynamics/ehrenfest_na.jl:23-49)."""
    from scipy.integrate import solve_ivp
    md = oracle.na_model(M=30)
    N, Ne, mass = 31, 15, 2000.0
    x0, v0, T, dt = 21.0, 0.0, 2000.0, 10.0
    occ0 = np.arange(Ne)[None, :]
    out = oracle.iesh_run_ensemble(1, md, Ne, mass, np.array([x0]), np.array([v0]), occ0, dt=dt, nsteps=int(T / dt))

    def rhs(t, y):
        x, v = y[0], y[1]
        psi = (y[2:2 + N * Ne] + 1j * y[2 + N * Ne:]).reshape(N, Ne)
        E, U, d = _na_nac(oracle, md, x)
        dh = md.mw2 * (x - md.g) - md.mw2 * x
        s = U[0] @ psi
        f = md.mw2 * x + dh * np.sum(np.abs(s) ** 2)
        dpsi = -1j * (E[:, None] * psi) - v * (d @ psi)
        return np.concatenate([[v, -f / mass], dpsi.real.ravel(), dpsi.imag.ravel()])

    psi0 = np.zeros((N, Ne), complex); psi0[np.arange(Ne), np.arange(Ne)] = 1
    y0 = np.concatenate([[x0, v0], psi0.real.ravel(), psi0.imag.ravel()])
    sol = solve_ivp(rhs, (0.0, T), y0, method="DOP853", rtol=1e-11, atol=1e-12)
    yT = sol.y[:, -1]
    assert np.isclose(out["r"][0] - x0, yT[0] - x0, rtol=1e-3)      # displacement ~0.075
    assert np.isclose(out["v"][0], yT[1], rtol=1e-3)
    psiT = (yT[2:2 + N * Ne] + 1j * yT[2 + N * Ne:]).reshape(N, Ne)
    assert np.allclose(out["psi"][0], psiT, rtol=1e-2, atol=1e-2)


def test_iesh_diabatic_population(oracle, key):
    """iesh_diabatic_population (iesh.jl:351-369): for psi = the occupied adiabatic orbitals it is the
    diabatic density sum_e U[a, s_e]^2 of the Slater determinant; it always sums to n_electrons."""
    md = oracle.na_model(M=20)
    N, Ne, mass = 21, 10, 2000.0
    r0 = np.array([7.0, 9.0]); v0 = np.array([5e-3, -5e-3])
    occ0 = np.tile(np.arange(Ne), (2, 1)); occ0[1, 0] = 15
    out = oracle.iesh_run_ensemble(0, md, Ne, mass, r0, v0, occ0, dt=2.0, nsteps=40, key=key, save_every=10)
    for i in range(2):
        E, U, _ = oracle.na_eigen(md, r0[i])
        assert np.allclose(out["pop_di"][i, 0], np.sum(U[:, occ0[i]] ** 2, axis=1), atol=1e-13)
    assert np.allclose(out["pop_di"].sum(axis=2), Ne, atol=1e-10)
    assert np.all(np.abs(out["pop_di"][:, -1] - out["pop_di"][:, 0]) > 0)


def _exact_chain(E, Ne, avail, beta, nmoves):
    """Distribution of the sorted occupied set after `nmoves` moves of the Boltzmann-factor chain
    (DynamicsUtils.jl:191-205), by enumerating its transitions: a uniform slot, a uniform free
    available state, acceptance min(1, exp(-beta dE))."""
    from collections import defaultdict
    dist = {tuple(avail[:Ne]): 1.0}
    for _ in range(nmoves):
        new = defaultdict(float)
        for st, p in dist.items():
            free = [a for a in avail if a not in st]
            for ci in range(Ne):
                for j in free:
                    x = -beta * (E[j] - E[st[ci]])
                    acc = 1.0 if x >= 0 else float(np.exp(x))
                    w = p / (Ne * len(free))
                    moved = list(st)
                    moved[ci] = j
                    new[tuple(moved)] += w * acc
                    new[st] += w * (1.0 - acc)
        dist = new
    out = defaultdict(float)
    for st, p in dist.items():
        out[tuple(sorted(st))] += p
    return out


@pytest.mark.parametrize("avail", [[0, 1, 2, 3, 4], [0, 2, 3, 4]])
def test_fermi_dirac_chain_distribution_is_exact(oracle, avail):
    """Empirical distribution of the oracle's occupations over 40000 Philox streams against the exact
    distribution of the reference's chain after n_avail * n_e moves (every transition enumerated):
    pins the initial state (first n_e available), the uniform choices, the acceptance rule and the
    restriction to the available states."""
    from collections import Counter
    E = np.array([-0.3, -0.1, 0.0, 0.15, 0.4])
    Ne, beta, n = 2, 4.0, 40000
    exact = _exact_chain(E, Ne, avail, beta, len(avail) * Ne)
    emp = Counter(tuple(oracle.fermi_dirac_occupations(E, Ne, beta, t, 12345, avail)) for t in range(n))
    assert set(emp) <= set(exact)
    for st, p in exact.items():
        assert abs(emp[st] / n - p) <= 5 * np.sqrt(p * (1 - p) / n) + 1e-4, st


def test_fermi_dirac_chain_length(oracle):
    """beta = 0 accepts every move: one electron on three states then walks the triangle, and after
    K moves sits on its start with probability (1 + 2 (-1/2)^K) / 3 -- 1/4 for the chain's
    n_avail * n_e = 3 moves (1/2 for 2, 3/8 for 4)."""
    E = np.zeros(3)
    n = 20000
    at_start = np.mean([oracle.fermi_dirac_occupations(E, 1, 0.0, t, 7)[0] == 0 for t in range(n)])
    assert abs(at_start - 0.25) < 5 * np.sqrt(0.25 * 0.75 / n)


def test_fermi_dirac_zero_temperature(oracle):
    """0 K leaves the lowest n_e states occupied (test/Dynamics/iesh.jl:37-41); equal energies give a
    NaN Boltzmann factor at beta = inf, which is never accepted (IEEE, as in the reference)."""
    md = oracle.na_model(M=30)
    E, _, _ = oracle.na_eigen(md, 1.5)
    for t in range(50):
        assert np.array_equal(oracle.fermi_dirac_occupations(E, 15, np.inf, t, 3), np.arange(15))
    avail = [1, 3, 4, 6]
    assert np.array_equal(oracle.fermi_dirac_occupations(np.zeros(8), 2, np.inf, 0, 3, avail), [1, 3])
    with pytest.raises(ValueError):
        oracle.fermi_dirac_occupations(E, 15, 1.0, 0, 3, np.arange(15))


def test_fermi_dirac_thermal_occupations(oracle):
    """Mean occupations of 5000 samples at 300 K against the Fermi function at the chemical potential
    that holds n_e electrons, normwise within the reference's rtol = 0.2 (test/Dynamics/iesh.jl:46-63;
    the fixed-n_e chain is canonical, so it is sharper than the grand-canonical function near mu)."""
    from scipy.optimize import brentq
    md = oracle.na_model(M=30)
    E, _, _ = oracle.na_eigen(md, 1.5)
    beta = 1.0 / (3.166811563e-6 * 300.0)
    n = 5000
    avg = np.zeros(31)
    for t in range(n):
        avg[oracle.fermi_dirac_occupations(E, 15, beta, t, 99)] += 1
    avg /= n
    fermi = lambda mu: 1.0 / (1.0 + np.exp(np.clip(beta * (E - mu), -700, 700)))
    mu = brentq(lambda m: fermi(m).sum() - 15, E.min() - 1, E.max() + 1)
    assert np.linalg.norm(avg - fermi(mu)) <= 0.2 * np.linalg.norm(fermi(mu))


def test_determinant_lower_bound_of_the_occupied_block():
    """The bound the device uses to skip the LU of the hop estimate: for orthonormal orbitals psi
    (N x n_e) and any n_e occupied rows, S = psi[occ, :] and T = psi[unoccupied, :] satisfy
    S^H S + T^H T = I, so |det S|^2 = prod(1 - s_i(T)^2) >= 1 - ||T||_F^2; hence the estimate
    2 dt sumvd (|Re det| + |Im det|) / |det|^2 <= 2 dt sumvd sqrt(2) / sqrt(1 - ||T||_F^2).  Checked on
    random isometries near and far from the occupied rows, including T nearly of norm 1."""
    rng = np.random.default_rng(17)
    for N, Ne in ((5, 2), (41, 20), (101, 50)):
        for scale in (1e-3, 0.1, 0.3, 1.0):
            X = np.eye(N, Ne) + scale * (rng.normal(size=(N, Ne)) + 1j * rng.normal(size=(N, Ne)))
            psi, _ = np.linalg.qr(X)
            occ = np.sort(rng.choice(N, Ne, replace=False)) if scale == 1.0 else np.arange(Ne)
            mask = np.zeros(N, bool); mask[occ] = True
            S, T = psi[mask], psi[~mask]
            d = np.linalg.det(S)
            t2 = np.sum(np.abs(T) ** 2)
            assert np.allclose(S.conj().T @ S + T.conj().T @ T, np.eye(Ne), atol=1e-12)
            assert abs(d) ** 2 >= 1 - t2 - 1e-12
            sumvd, dt = 0.37, 2.0
            if 1 - t2 > 0:
                est = 2 * dt / abs(d) ** 2 * sumvd * (abs(d.real) + abs(d.imag))
                assert est <= 2 * dt * sumvd * np.sqrt(2) / np.sqrt(1 - t2) * (1 + 1e-12)


def _na_gauge_fixed(oracle, md, x):
    """Diagonal of the model matrix and the eigenpairs with the impurity component of every
    eigenvector positive (the device's gauge)."""
    E, U, _ = oracle.na_eigen(md, x)
    U = U * np.sign(U[0])
    return np.diag(oracle.na_matrix(md, x)).copy(), E, U


def test_fermi_dirac_diabatic_orbitals(oracle):
    """FermiDiracState{Diabatic} (iesh.jl:138-183): every orbital is a row of the eigenvector matrix
    (the diabatic state in the adiabatic basis) with the sign that makes its first adiabatic
    component positive; the orbitals are orthonormal and the adiabatic occupations distinct and
    sorted."""
    md = oracle.na_model(M=20)
    Ed, E, U = _na_gauge_fixed(oracle, md, 3.0)
    assert np.isclose(Ed[0], 0.5 * 2000 * 2e-4 ** 2 * ((3.0 - 20.6097) ** 2 - 9.0) - 3.8e-3)
    for t in range(20):
        occ, psi = oracle.fermi_dirac_diabatic(Ed, U, 10, 1.0 / (3.166811563e-6 * 300), t, 5)
        assert len(set(occ)) == 10 and np.all(np.diff(occ) > 0)
        assert np.allclose(psi.T @ psi, np.eye(10), atol=1e-12)
        for i in range(10):
            d = int(np.argmax(np.abs(U @ psi[:, i])))  # U U^T e_d = e_d
            assert np.allclose(psi[:, i], U[d] * np.sign(U[d, 0]), atol=1e-14)
            assert psi[0, i] > 0


def test_fermi_dirac_diabatic_adiabatic_state_distribution(oracle):
    """The adiabatic states drawn for fixed diabatic states (0 K and one higher available level, so the
    chain keeps its start): one electron on bath level 5 lands on adiabatic state s with probability
    U[5, s]^2; two electrons on levels 3 and 9 give the sorted pair {a, b} with probability
    w3_a w9_b / (1 - w9_a) + w3_b w9_a / (1 - w9_b) (the second draw repeats until it misses the first)."""
    from collections import Counter
    md = oracle.na_model(M=20)
    Ed, E, U = _na_gauge_fixed(oracle, md, 3.0)
    n = 20000
    w = U[6] ** 2  # bath level 5 is state 6
    cnt = Counter(int(oracle.fermi_dirac_diabatic(Ed, U, 1, np.inf, t, 9, [6, 16])[0][0]) for t in range(n))
    for s in range(21):
        assert abs(cnt[s] / n - w[s]) <= 5 * np.sqrt(w[s] * (1 - w[s]) / n) + 1e-3
    w1, w2 = U[4] ** 2, U[10] ** 2
    cnt = Counter(tuple(oracle.fermi_dirac_diabatic(Ed, U, 2, np.inf, t, 9, [4, 10, 18])[0]) for t in range(n))
    for a in range(21):
        for b in range(a + 1, 21):
            p = w1[a] * w2[b] / (1 - w2[a]) + w1[b] * w2[a] / (1 - w2[b])
            assert abs(cnt[(a, b)] / n - p) <= 5 * np.sqrt(p * (1 - p) / n) + 1e-3, (a, b)


def test_ring_polymer_iesh_conserves_ring_polymer_energy(oracle):
    """BCBWavefunction for RingPolymerSimulation{AdiabaticIESH} (bcb_wavefunction.jl, rpiesh.jl:20-38)
    without hops: the beads move on U0 + sum_occ E_k with the ring-polymer springs, and the ring-polymer
    energy sum_b [m v_b^2 / 2 + U0(x_b) + sum_occ E_k(x_b)] + m w_n^2 / 2 sum_b (x_b - x_b+1)^2 is kept
    to 1e-9 with a second-order error (a wrong force sign, mode frequency or Cayley entry breaks it);
    EhrenfestNA keeps every electron normalised."""
    M, Ne, P, mass = 20, 10, 4, 2000.0
    md = oracle.na_model(M=M)
    mw2, temp = 2000 * 2e-4 ** 2, 3.166811563e-6 * 300.0
    wn = P * temp
    rng = np.random.default_rng(3)
    r0 = rng.normal(-3.0, 0.3, (8, P))
    v0 = rng.normal(0.0, 2e-3, (8, P))
    occ0 = np.tile(np.arange(Ne), (8, 1))

    def energy(r, v, occ):
        e = 0.5 * mass * np.sum(v ** 2) + 0.5 * mass * wn ** 2 * np.sum((r - np.roll(r, -1)) ** 2)
        for b in range(P):
            e += 0.5 * mw2 * r[b] ** 2 + oracle.na_eigen(md, r[b])[0][occ].sum()
        return e

    errs = []
    for dt in (1.0, 0.5):
        c = oracle.rp_iesh_run_ensemble(0, md, Ne, mass, P, temp, r0, v0, occ0, dt=dt, nsteps=int(1000 / dt), key=5)
        assert c["nhops"].sum() == 0
        errs.append(max(abs(energy(c["r"][i], c["v"][i], c["occ"][i]) - energy(r0[i], v0[i], occ0[i])) for i in range(8)))
    assert errs[0] < 1e-9 and 3.0 < errs[0] / errs[1] < 5.0
    c = oracle.rp_iesh_run_ensemble(1, md, Ne, mass, P, temp, r0, v0, occ0, dt=1.0, nsteps=300, key=5)
    assert np.allclose(np.sum(np.abs(c["psi"]) ** 2, axis=1), 1.0, atol=1e-11)
