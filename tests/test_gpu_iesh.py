"""GPU IESH / EhrenfestNA (Newns-Anderson, VerletwithElectronics) against the FP64 oracle:
identical occupations and hop counts, positions, velocities, wavefunctions and impurity
populations within 1e-9 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MASS = 2000.0
MODEL = dict(m=2000.0, omega=2e-4, g=20.6097, dG=-3.8e-3, Gamma=6.4e-3)


def _params(M):
    W = 6 * MODEL["Gamma"] / 2
    return [MODEL["m"], MODEL["omega"], MODEL["g"], MODEL["dG"], MODEL["Gamma"], -W, W, M]


def _inputs(ntraj, Ne, seed, x0=8.0):
    rng = np.random.default_rng(seed)
    r0 = rng.normal(x0, 1.0, size=ntraj)
    v0 = np.abs(rng.normal(0.0, 1e-2, size=ntraj))
    occ0 = np.tile(np.arange(Ne), (ntraj, 1)).astype(np.int32)
    occ0[:, -1] = Ne  # one electron above the Fermi level
    return r0, v0, occ0


def _gpu(L, method, M, Ne, r0, v0, occ0, dt, nsteps, key, save_every=10, psi=None, **kw):
    E = L.Ensemble(0)
    E.set_problem(method, M + 1, 1, 1, [MASS], dt, algorithm=L.ALG_VERLET_ELECTRONICS, n_electrons=Ne, **kw)
    E.set_model(0, L.MODEL_NEWNS_ANDERSON, _params(M))
    E.load_initial_iesh(r0, v0, occ0, psi)
    stats = E.run(0.0, nsteps * dt, key, save_every=save_every)
    out = {"r": E.download(L.FIELD_R), "v": E.download(L.FIELD_V),
           "psi": E.download(L.FIELD_SIGMA_RE) + 1j * E.download(L.FIELD_SIGMA_IM),
           "occ": E.download(L.FIELD_STATE), "nhops": E.download(L.FIELD_NHOPS),
           "di": E.download(L.FIELD_SAVE_POP_DI), "ad": E.download(L.FIELD_SAVE_POP_AD),
           "di_mean": E.reduce(L.OUT_DIABATIC_POPULATION, L.RED_MEAN, E.nsave * (M + 1)).reshape(E.nsave, M + 1),
           "stats": stats}
    E.close()
    return out


def _cpu(oracle, method, M, Ne, r0, v0, occ0, dt, nsteps, key, save_every=10, **kw):
    W = 6 * MODEL["Gamma"] / 2
    md = oracle.na_model(M=M, W=W, m=MODEL["m"], omega=MODEL["omega"], g=MODEL["g"], dG=MODEL["dG"], Gamma=MODEL["Gamma"])
    return oracle.iesh_run_ensemble(method, md, Ne, MASS, r0, v0, occ0, dt=dt, nsteps=nsteps, key=key,
                                    save_every=save_every, **kw)


def _compare(g, c, hops=True):
    if hops:
        assert np.array_equal(g["occ"], c["occ"])
        assert np.array_equal(g["nhops"], c["nhops"])
    assert np.allclose(g["r"], c["r"], rtol=1e-9, atol=1e-11)
    assert np.allclose(g["v"], c["v"], rtol=1e-9, atol=1e-13)
    assert np.allclose(g["psi"], c["psi"], rtol=1e-9, atol=1e-10)
    assert np.allclose(g["di"], c["pop_di"], rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("M,Ne,ntraj,nsteps", [(40, 20, 64, 300), (100, 50, 12, 120)])
def test_iesh_fp64_parity(oracle, key, M, Ne, ntraj, nsteps):
    from nqcd_mi355x import _lib as L
    r0, v0, occ0 = _inputs(ntraj, Ne, 11)
    g = _gpu(L, L.METHOD_IESH, M, Ne, r0, v0, occ0, 2.0, nsteps, key)
    c = _cpu(oracle, 0, M, Ne, r0, v0, occ0, 2.0, nsteps, key)
    if M == 40:
        assert c["nhops"].sum() >= 3
    _compare(g, c)
    # one-hot adiabatic populations at the last save slot: a save holds the state before the step's hop
    # callback, so the slot of the last step shows the occupations after nsteps - 1 steps
    prev = _cpu(oracle, 0, M, Ne, r0, v0, occ0, 2.0, nsteps - 1, key)
    last = np.zeros((ntraj, M + 1))
    np.put_along_axis(last, prev["occ"], 1.0, axis=1)
    assert np.array_equal(g["ad"][:, -1, :], last)
    assert np.allclose(g["di_mean"], c["pop_di"].mean(axis=0), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("M,Ne,ntraj,nsteps", [(2, 1, 32, 200), (16, 8, 32, 200), (40, 30, 16, 120), (63, 40, 8, 80),
                                              (127, 20, 8, 60)])
def test_iesh_fp64_parity_sizes(oracle, key, M, Ne, ntraj, nsteps):
    """The smallest problem (3 states, one electron, one wavefront), a two-wavefront block, the wide
    column-tile variant (2 n_e above the padded state count: N = 41 with 30 electrons, N = 64 with
    40) and the largest state count (N = 128)."""
    from nqcd_mi355x import _lib as L
    r0, v0, occ0 = _inputs(ntraj, Ne, 21)
    g = _gpu(L, L.METHOD_IESH, M, Ne, r0, v0, occ0, 2.0, nsteps, key)
    c = _cpu(oracle, 0, M, Ne, r0, v0, occ0, 2.0, nsteps, key)
    _compare(g, c)


def test_iesh_too_large_for_lds_is_rejected(key):
    """128 states with 64 electrons would need more than a compute unit's 160 KB of LDS: the run
    fails with a message naming the sizes instead of launching."""
    from nqcd_mi355x import _lib as L
    r0, v0, occ0 = _inputs(4, 64, 22)
    with pytest.raises(L.NQCDError, match="LDS"):
        _gpu(L, L.METHOD_IESH, 127, 64, r0, v0, occ0, 2.0, 10, key)


def test_iesh_estimate_off_and_rescaling_modes(oracle, key):
    from nqcd_mi355x import _lib as L
    M, Ne, ntraj, nsteps = 40, 20, 32, 200
    r0, v0, occ0 = _inputs(ntraj, Ne, 12)
    for resc, ores in ((L.RESCALE_VINVERSION, 1), (L.RESCALE_OFF, 2)):
        g = _gpu(L, L.METHOD_IESH, M, Ne, r0, v0, occ0, 2.0, nsteps, key, rescaling=resc, estimate_probability=False)
        c = _cpu(oracle, 0, M, Ne, r0, v0, occ0, 2.0, nsteps, key, rescaling=ores, estimate=False)
        _compare(g, c)


def test_iesh_edc_parity(oracle, key):
    from nqcd_mi355x import _lib as L
    M, Ne, ntraj, nsteps = 40, 20, 32, 200
    r0, v0, occ0 = _inputs(ntraj, Ne, 13)
    g = _gpu(L, L.METHOD_IESH, M, Ne, r0, v0, occ0, 2.0, nsteps, key, decoherence=L.DECOHERENCE_EDC, edc_C=0.1)
    c = _cpu(oracle, 0, M, Ne, r0, v0, occ0, 2.0, nsteps, key, edc=True, edc_C=0.1)
    _compare(g, c)
    assert np.allclose(np.sum(np.abs(g["psi"]) ** 2, axis=1), 1.0, atol=1e-12)


@pytest.mark.parametrize("M,Ne", [(40, 20), (100, 50)])
def test_ehrenfest_na_fp64_parity(oracle, key, M, Ne):
    from nqcd_mi355x import _lib as L
    ntraj, nsteps = 16, 150
    r0, v0, occ0 = _inputs(ntraj, Ne, 14)
    g = _gpu(L, L.METHOD_EHRENFEST_NA, M, Ne, r0, v0, occ0, 2.0, nsteps, key)
    c = _cpu(oracle, 1, M, Ne, r0, v0, occ0, 2.0, nsteps, key)
    _compare(g, c, hops=False)
    # adiabatic population sum_e |psi_ae|^2 sums to the electron count
    assert np.allclose(g["ad"].sum(axis=2), Ne, atol=1e-10)


def test_iesh_abi_errors():
    from nqcd_mi355x import _lib as L
    E = L.Ensemble(0)
    with pytest.raises(L.NQCDError):  # wrong algorithm
        E.set_problem(L.METHOD_IESH, 41, 1, 1, [MASS], 1.0, algorithm=L.ALG_BAB_TSIT5, n_electrons=20)
    with pytest.raises(L.NQCDError):  # too many electrons
        E.set_problem(L.METHOD_IESH, 41, 1, 1, [MASS], 1.0, algorithm=L.ALG_VERLET_ELECTRONICS, n_electrons=41)
    E.set_problem(L.METHOD_IESH, 41, 1, 1, [MASS], 1.0, algorithm=L.ALG_VERLET_ELECTRONICS, n_electrons=20)
    E.set_model(0, L.MODEL_NEWNS_ANDERSON, _params(40))
    occ = np.tile(np.arange(20), (2, 1))
    occ[1, 3] = 2  # state 2 occupied twice
    with pytest.raises(L.NQCDError):
        E.load_initial_iesh([8.0, 8.0], [0.0, 0.0], occ)
    E.close()


def test_device_determinant_matches_numpy():
    """The single-wavefront LU determinant behind the hopping estimate: random complex matrices of
    every size up to 64, permutation matrices (the sign of the pivot permutation), matrices with
    tied pivot candidates, a singular matrix, and the occupied-row submatrices of unitary matrices
    (the IESH case, |det| <= 1)."""
    from nqcd_mi355x import _lib as L
    rng = np.random.default_rng(3)
    E = L.Ensemble(0)
    try:
        for n in list(range(1, 17)) + [20, 31, 32, 33, 50, 63, 64]:
            S = rng.normal(size=(6, n, n)) + 1j * rng.normal(size=(6, n, n))
            d, ok = E.test_determinant(S)
            ref = np.linalg.det(S)
            assert ok.all()
            assert np.all(np.abs(d - ref) <= 1e-12 * np.abs(ref)), n
        for n in (2, 3, 5, 8, 17, 64):
            P = np.stack([np.eye(n)[rng.permutation(n)] for _ in range(4)]).astype(complex)
            d, ok = E.test_determinant(P)
            assert ok.all() and np.allclose(d, np.linalg.det(P), atol=0)
        T = np.ones((1, 4, 4), complex) + np.diag([0.0, 1.0, 2.0, 3.0])[None]  # ties in the first column
        d, _ = E.test_determinant(T)
        assert abs(d[0] - np.linalg.det(T[0])) < 1e-12
        Z = rng.normal(size=(1, 5, 5)) + 0j
        Z[0, :, 2] = 0.0
        d, ok = E.test_determinant(Z)
        assert not ok[0] and d[0] == 0
        for (N, Ne) in ((41, 20), (101, 50)):
            Q, _ = np.linalg.qr(rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
            occ = np.sort(rng.choice(N, Ne, replace=False))
            S = Q[occ][:, :Ne][None]
            d, ok = E.test_determinant(S)
            ref = np.linalg.det(S[0])
            assert ok[0] and abs(d[0] - ref) <= 1e-11 * abs(ref)
    finally:
        E.close()


def _fd_occupations(L, method, M, Ne, r0, beta, seed, avail=None, nsteps=0, key=0, diabatic=False):
    """Device Fermi-Dirac occupations at r0 (and, with nsteps > 0, the run that follows them)."""
    E = L.Ensemble(0)
    try:
        E.set_problem(method, M + 1, 1, 1, [MASS], 2.0, algorithm=L.ALG_VERLET_ELECTRONICS, n_electrons=Ne)
        E.set_model(0, L.MODEL_NEWNS_ANDERSON, _params(M))
        v0 = np.full(len(r0), 5e-3)
        E.load_initial_iesh(r0, v0, np.tile(np.arange(Ne), (len(r0), 1)))
        E.sample_fermi_dirac(beta, seed, avail, diabatic=diabatic)
        E.run(0.0, nsteps * 2.0, key, save_every=10 if nsteps else 0)
        out = {"occ": E.download(L.FIELD_STATE), "r": E.download(L.FIELD_R), "v": E.download(L.FIELD_V),
               "psi": E.download(L.FIELD_SIGMA_RE) + 1j * E.download(L.FIELD_SIGMA_IM),
               "nhops": E.download(L.FIELD_NHOPS), "di": E.download(L.FIELD_SAVE_POP_DI)}
    finally:
        E.close()
    return out, v0


def _fd_oracle(oracle, M, Ne, r0, beta, seed, avail=None):
    W = 6 * MODEL["Gamma"] / 2
    md = oracle.na_model(M=M, W=W, m=MODEL["m"], omega=MODEL["omega"], g=MODEL["g"], dG=MODEL["dG"], Gamma=MODEL["Gamma"])
    av = None if avail is None else sorted(avail)
    return np.stack([oracle.fermi_dirac_occupations(oracle.na_eigen(md, x)[0], Ne, beta, t, seed, av)
                     for t, x in enumerate(r0)])


KT300 = 3.166811563e-6 * 300.0


@pytest.mark.parametrize("M,Ne,beta,avail", [(40, 20, 1 / KT300, None), (100, 50, 1 / KT300, None),
                                             (4, 2, 0.0, None), (30, 15, 1 / (4 * KT300), [0, 2, 5] + list(range(9, 31))),
                                             (127, 20, 0.0, None), (127, 20, 1 / KT300, None)])
def test_fermi_dirac_occupations_match_oracle(oracle, M, Ne, beta, avail):
    """Device chain (one lane per trajectory, the adiabatic energies from the secular roots) against the
    oracle's literal restatement of sample_fermi_dirac_distribution on the same Philox streams:
    identical occupations at 300 K, at infinite temperature, with a subset of available states and at
    the largest state count (two 64-bit occupancy words; at infinite temperature electrons reach the
    upper word)."""
    from nqcd_mi355x import _lib as L
    rng = np.random.default_rng(M)
    r0 = rng.normal(3.0, 2.0, size=96)
    g, _ = _fd_occupations(L, L.METHOD_IESH, M, Ne, r0, beta, 0x5EED, avail)
    c = _fd_oracle(oracle, M, Ne, r0, beta, 0x5EED, avail)
    assert np.array_equal(g["occ"], c)
    if beta > 0:
        assert not np.all(c == np.arange(Ne))  # thermal: some trajectories leave the ground state


def test_fermi_dirac_zero_temperature_and_errors():
    """0 K gives the lowest n_e states (test/Dynamics/iesh.jl:37-41); too few available states and an
    FSSH problem are rejected."""
    from nqcd_mi355x import _lib as L
    g, _ = _fd_occupations(L, L.METHOD_IESH, 40, 20, np.linspace(-5, 25, 64), np.inf, 1)
    assert np.array_equal(g["occ"], np.tile(np.arange(20), (64, 1)))
    with pytest.raises(L.NQCDError, match="available"):
        _fd_occupations(L, L.METHOD_IESH, 40, 20, np.zeros(4), 1.0, 1, avail=list(range(20)))
    E = L.Ensemble(0)
    try:
        E.set_problem(L.METHOD_FSSH, 2, 1, 1, [MASS], 1.0)
        with pytest.raises(L.NQCDError, match="IESH"):
            E.sample_fermi_dirac(1.0, 1)
    finally:
        E.close()


@pytest.mark.parametrize("method", [0, 1])
def test_fermi_dirac_then_dynamics_parity(oracle, key, method):
    """The whole DynamicsVariables(sim, v, r, FermiDiracState) -> run path: device-sampled thermal
    occupations, then IESH / EhrenfestNA from them, against the oracle run from the oracle's
    occupations (1e-9 relative, identical hops)."""
    from nqcd_mi355x import _lib as L
    M, Ne, nsteps = 40, 20, 150
    r0 = np.random.default_rng(5).normal(6.0, 1.0, size=32)
    meth = L.METHOD_IESH if method == 0 else L.METHOD_EHRENFEST_NA
    g, v0 = _fd_occupations(L, meth, M, Ne, r0, 1 / KT300, 77, nsteps=nsteps, key=key)
    occ0 = _fd_oracle(oracle, M, Ne, r0, 1 / KT300, 77)
    c = _cpu(oracle, method, M, Ne, r0, v0, occ0, 2.0, nsteps, key)
    _compare(g, c, hops=method == 0)


@pytest.mark.parametrize("orthonormal", [True, False])
def test_iesh_given_wavefunction_parity(oracle, key, orthonormal):
    """A caller-given psi (load_initial with sig_re / sig_im): orthonormal orbitals mixed away from the
    occupied states, where the device may decide the hop estimate from the bound |det S|^2 >= 1 - t2
    without the LU, and orbitals 5% off normalisation, where it must not; both against the oracle
    (which always takes the LU), identical hops."""
    from nqcd_mi355x import _lib as L
    M, Ne, ntraj, nsteps = 40, 20, 32, 200
    r0, v0, occ0 = _inputs(ntraj, Ne, 15)
    rng = np.random.default_rng(16)
    psi = np.zeros((ntraj, M + 1, Ne), complex)
    for i in range(ntraj):
        X = np.zeros((M + 1, Ne), complex)
        X[occ0[i], np.arange(Ne)] = 1.0
        X += 0.02 * (rng.normal(size=X.shape) + 1j * rng.normal(size=X.shape))  # t2 ~ 0.3
        psi[i] = np.linalg.qr(X)[0]
        if not orthonormal:
            psi[i] *= 1.05
    g = _gpu(L, L.METHOD_IESH, M, Ne, r0, v0, occ0, 2.0, nsteps, key, psi=psi)
    c = _cpu(oracle, 0, M, Ne, r0, v0, occ0, 2.0, nsteps, key, psi0=psi)
    assert c["nhops"].sum() >= 1
    _compare(g, c)


def _fd_diabatic_oracle(oracle, M, Ne, r0, beta, seed, avail=None):
    W = 6 * MODEL["Gamma"] / 2
    md = oracle.na_model(M=M, W=W, m=MODEL["m"], omega=MODEL["omega"], g=MODEL["g"], dG=MODEL["dG"], Gamma=MODEL["Gamma"])
    occ, psi = [], []
    for t, x in enumerate(r0):
        E, U, _ = oracle.na_eigen(md, x)
        U = U * np.sign(U[0])  # impurity component positive, as on the device
        o, p = oracle.fermi_dirac_diabatic(np.diag(oracle.na_matrix(md, x)), U, Ne, beta, t, seed, avail)
        occ.append(o); psi.append(p)
    return np.stack(occ), np.stack(psi)


@pytest.mark.parametrize("M,Ne,beta,avail", [(40, 20, 1 / KT300, None), (100, 50, 1 / KT300, None),
                                             (30, 15, 1 / (4 * KT300), [0, 2, 5] + list(range(9, 31))),
                                             (127, 20, 1 / KT300, None)])
def test_fermi_dirac_diabatic_matches_oracle(oracle, M, Ne, beta, avail):
    """FermiDiracState{Diabatic} (iesh.jl:138-183) on the device -- the chain over the model matrix's
    diagonal, the weighted draws of adiabatic states, orbitals from rows of the closed-form
    eigenvectors -- against the oracle's literal restatement (Jacobi eigenvectors in the same gauge):
    identical occupations, orbitals within 1e-12."""
    from nqcd_mi355x import _lib as L
    r0 = np.random.default_rng(M + 1).normal(3.0, 2.0, size=64)
    g, _ = _fd_occupations(L, L.METHOD_IESH, M, Ne, r0, beta, 0xD1AB, avail, diabatic=True)
    occ, psi = _fd_diabatic_oracle(oracle, M, Ne, r0, beta, 0xD1AB, avail)
    assert np.array_equal(g["occ"], occ)
    assert np.allclose(g["psi"], psi, atol=1e-12)


def test_fermi_dirac_diabatic_then_iesh_parity(oracle, key):
    """Diabatic Fermi-Dirac initial orbitals (device) followed by IESH, against the oracle run from the
    oracle's occupations and orbitals: identical hops, 1e-9 relative."""
    from nqcd_mi355x import _lib as L
    M, Ne, nsteps = 40, 20, 150
    r0 = np.random.default_rng(6).normal(6.0, 1.0, size=32)
    g, v0 = _fd_occupations(L, L.METHOD_IESH, M, Ne, r0, 1 / KT300, 78, nsteps=nsteps, key=key, diabatic=True)
    occ, psi = _fd_diabatic_oracle(oracle, M, Ne, r0, 1 / KT300, 78)
    c = _cpu(oracle, 0, M, Ne, r0, v0, occ, 2.0, nsteps, key, psi0=psi.astype(complex))
    _compare(g, c)


def test_fermi_dirac_diabatic_is_iesh_only():
    from nqcd_mi355x import _lib as L
    with pytest.raises(L.NQCDError, match="IESH"):
        _fd_occupations(L, L.METHOD_EHRENFEST_NA, 40, 20, np.zeros(4), 1.0, 1, diabatic=True)


def _rp_inputs(ntraj, Ne, P, seed):
    r0, v0, occ0 = _inputs(ntraj, Ne, seed)
    rng = np.random.default_rng(seed + 1)
    return r0[:, None] + rng.normal(0, 0.05, (ntraj, P)), v0[:, None] + rng.normal(0, 1e-3, (ntraj, P)), occ0


@pytest.mark.parametrize("method,M,Ne,P,ntraj,nsteps,resc", [(0, 40, 20, 4, 32, 200, 0), (0, 20, 10, 8, 32, 200, 1),
                                                             (1, 40, 20, 4, 16, 150, 0), (0, 63, 30, 2, 6, 80, 0)])
def test_ring_polymer_iesh_parity(oracle, key, method, M, Ne, P, ntraj, nsteps, resc):
    """RingPolymerSimulation{AdiabaticIESH / EhrenfestNA} with BCBWavefunction (bcb_wavefunction.jl:37-69):
    bead forces from per-bead eigenpairs, Cayley free ring-polymer step in normal modes, the
    wavefunction step at the centroid of the step start, hops on the centroid with every bead
    rescaled; against the oracle: identical hops, beads, psi and populations within 1e-9."""
    from nqcd_mi355x import _lib as L
    temp = 3.166811563e-6 * 300.0
    rb, vb, occ0 = _rp_inputs(ntraj, Ne, P, 31)
    E = L.Ensemble(0)
    try:
        E.set_problem(L.METHOD_IESH if method == 0 else L.METHOD_EHRENFEST_NA, M + 1, 1, 1, [MASS], 2.0,
                      algorithm=L.ALG_BCB_WAVEFUNCTION, nbeads=P, temperature=temp, n_electrons=Ne, rescaling=resc)
        E.set_model(0, L.MODEL_NEWNS_ANDERSON, _params(M))
        E.load_initial_iesh(rb, vb, occ0)
        E.run(0.0, nsteps * 2.0, key, save_every=10)
        g = {"r": E.download(L.FIELD_R), "v": E.download(L.FIELD_V),
             "psi": E.download(L.FIELD_SIGMA_RE) + 1j * E.download(L.FIELD_SIGMA_IM),
             "occ": E.download(L.FIELD_STATE), "nhops": E.download(L.FIELD_NHOPS), "di": E.download(L.FIELD_SAVE_POP_DI)}
    finally:
        E.close()
    W = 6 * MODEL["Gamma"] / 2
    md = oracle.na_model(M=M, W=W, m=MODEL["m"], omega=MODEL["omega"], g=MODEL["g"], dG=MODEL["dG"], Gamma=MODEL["Gamma"])
    c = oracle.rp_iesh_run_ensemble(method, md, Ne, MASS, P, temp, rb, vb, occ0, dt=2.0, nsteps=nsteps, key=key,
                                    rescaling=resc, save_every=10)
    if method == 0:
        assert np.array_equal(g["occ"], c["occ"])
        assert np.array_equal(g["nhops"], c["nhops"])
        if M == 40:
            assert c["nhops"].sum() >= 3
    assert g["r"].shape == (ntraj, P)
    assert np.allclose(g["r"], c["r"], rtol=1e-9, atol=1e-11)
    assert np.allclose(g["v"], c["v"], rtol=1e-9, atol=1e-13)
    assert np.allclose(g["psi"], c["psi"], rtol=1e-9, atol=1e-10)
    assert np.allclose(g["di"], c["pop_di"], rtol=1e-9, atol=1e-11)


def test_ring_polymer_iesh_rejects_edc_and_verlet():
    from nqcd_mi355x import _lib as L
    E = L.Ensemble(0)
    try:
        with pytest.raises(L.NQCDError, match="BCBWavefunction"):
            E.set_problem(L.METHOD_IESH, 41, 1, 1, [MASS], 2.0, algorithm=L.ALG_VERLET_ELECTRONICS, nbeads=4,
                          temperature=1e-3, n_electrons=20)
        with pytest.raises(L.NQCDError, match="EDC"):
            E.set_problem(L.METHOD_IESH, 41, 1, 1, [MASS], 2.0, algorithm=L.ALG_BCB_WAVEFUNCTION, nbeads=4,
                          temperature=1e-3, n_electrons=20, decoherence=L.DECOHERENCE_EDC)
    finally:
        E.close()
