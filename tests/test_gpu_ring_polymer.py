"""GPU ring-polymer dynamics (RPMD with BCB; RPSH and RP-Ehrenfest with BCBwithTsit5) against
the FP64 oracle, plus the analytic harmonic ring polymer."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
MASS = 1836.0
BETA = 8.0
DW1 = [1 / 1200, 1 / 40, 0.0, 0.0, 1.0]     # single-state double well
DW2 = [1 / 1200, 1 / 40, 0.01, 0.005, 1.0]  # + second diabat offset 0.01, Gaussian coupling


def _ens(L, method, P, nstates, dt=1.0, precision=0, rescaling=0):
    E = L.Ensemble(0)
    alg = L.ALG_BCB if method == L.METHOD_RPMD else L.ALG_BCB_TSIT5
    E.set_problem(method, nstates, 1, 1, [MASS], dt, algorithm=alg, nbeads=P, precision=precision,
                  temperature=1.0 / BETA, rescaling=rescaling)
    E.set_model(0, L.MODEL_DW_GAUSSIAN, DW2 if nstates == 2 else DW1)
    return E


def _thermal_beads(L, P, ntraj, key, x0=-3.87, nstates=1):
    """Device sampler, read back through a zero-step run."""
    method = L.METHOD_RPMD if nstates == 1 else L.METHOD_RPSH
    E = _ens(L, method, P, nstates)
    E.sample_initial(L.SAMPLER_RING_POLYMER_THERMAL, ntraj, [x0, np.sqrt(0.1 / MASS), BETA], key)
    E.run(0.0, 0.0, key)
    r, v = E.download(L.FIELD_R), E.download(L.FIELD_V)
    E.close()
    return r, v


def _layout(monkeypatch, layout, P, exists):
    """Selects the ring-polymer layout: one ring polymer per lane or P lanes per ring polymer."""
    if layout == "lane" and not exists:
        pytest.skip(f"no one-ring-polymer-per-lane kernel at P = {P}")
    monkeypatch.setenv("NQCD_MI355X_RP_LAYOUT", layout)


@pytest.mark.parametrize("layout", ["bead", "lane"])
@pytest.mark.parametrize("P", [1, 2, 4, 8, 16, 32, 64])
def test_rpmd_fp64_parity(oracle, key, P, layout, monkeypatch):
    from nqcd_mi355x import _lib as L
    _layout(monkeypatch, layout, P, P <= 32)
    ntraj, nsteps = 256, 500
    r0, v0 = _thermal_beads(L, P, ntraj, key)
    assert np.isfinite(r0).all() and abs(r0.mean() + 3.87) < 0.2
    E = _ens(L, L.METHOD_RPMD, P, 1)
    E.load_initial(r0, v0, np.zeros((ntraj, 2, 2)))
    E.run(0.0, float(nsteps), key, save_every=50)
    g_r, g_v, g_c = E.download(L.FIELD_R), E.download(L.FIELD_V), E.download(L.FIELD_SAVE_CENTROID)
    E.close()
    m = oracle.make_model(oracle.MODEL_DW_GAUSSIAN, 1, 1, DW1)
    c = oracle.rp_run_ensemble(oracle.RP_CLASSICAL, m, MASS, r0, v0, dt=1.0, nsteps=nsteps, temperature=1 / BETA,
                               save_every=50, nsave=nsteps // 50 + 1, key=key)
    assert np.allclose(g_r, c["r"], rtol=1e-9, atol=1e-10)
    assert np.allclose(g_v, c["v"], rtol=1e-9, atol=1e-12)
    assert np.allclose(g_c, c["centroid"], rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("layout", ["bead", "lane"])
@pytest.mark.parametrize("P", [1, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("method", ["rpsh", "rpeh"])
def test_rp_electronic_fp64_parity(oracle, key, P, method, layout, monkeypatch):
    from nqcd_mi355x import _lib as L
    _layout(monkeypatch, layout, P, method == "rpsh" and P <= 32)
    ntraj, nsteps = 256, 600
    rng = np.random.default_rng(P)
    # ring polymers crossing the coupling region at x ~ 0 with a spread of velocities
    r0 = rng.normal(-2.0, 0.05, size=(ntraj, 1)) + rng.normal(0, 0.05, size=(ntraj, P))
    v0 = rng.uniform(0.011, 0.018, size=(ntraj, 1)) + rng.normal(0, 1e-4, size=(ntraj, P))
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    st0 = np.zeros(ntraj, np.int32)
    lm = L.METHOD_RPSH if method == "rpsh" else L.METHOD_RP_EHRENFEST
    E = _ens(L, lm, P, 2)
    E.load_initial(r0, v0, sig0, state=st0)
    E.run(0.0, float(nsteps), key, save_every=10)
    g = {"r": E.download(L.FIELD_R), "v": E.download(L.FIELD_V), "state": E.download(L.FIELD_STATE),
         "sig": E.download(L.FIELD_SIGMA_RE) + 1j * E.download(L.FIELD_SIGMA_IM),
         "save_state": E.download(L.FIELD_SAVE_STATE), "pop_di": E.download(L.FIELD_SAVE_POP_DI)}
    E.close()
    m = oracle.make_model(oracle.MODEL_DW_GAUSSIAN, 2, 1, DW2)
    om = oracle.RP_FSSH if method == "rpsh" else oracle.RP_EHRENFEST
    c = oracle.rp_run_ensemble(om, m, MASS, r0, v0, sig0, st0, dt=1.0, nsteps=nsteps, temperature=1 / BETA,
                               save_every=10, nsave=nsteps // 10 + 1, key=key)
    if method == "rpsh":
        assert c["nhops"].sum() > 10
        assert np.array_equal(g["save_state"], c["save_state"])
    assert np.allclose(g["r"], c["r"], rtol=1e-9, atol=1e-10)
    # bead velocities reach |v| ~ 1 through the stiff internal modes at large P, while some elements
    # cross zero: 1e-9 relative to the largest velocity (the observed differences are 5e-12 at P = 64)
    assert np.allclose(g["v"], c["v"], rtol=1e-9, atol=1e-9 * np.abs(c["v"]).max())
    assert np.allclose(g["sig"], c["sig_re"] + 1j * c["sig_im"], rtol=1e-9, atol=1e-10)
    assert np.allclose(g["pop_di"], c["pop_di"], rtol=1e-9, atol=1e-10)


def test_rpmd_harmonic_matches_analytic_normal_modes(oracle, key):
    """Harmonic ring polymer: normal mode k oscillates at sqrt(w_k^2 + w^2) (the reference's
    BCB convergence test, test/Dynamics/algorithms/bcb.jl:18-44); dt small so the error is ~dt^2."""
    from nqcd_mi355x import _lib as L
    P, T, w0, dt, tend = 16, 0.3, 1.0, 0.002, 5.0
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_RPMD, 1, 1, 1, [1.0], dt, algorithm=L.ALG_BCB, nbeads=P, temperature=T)
    E.set_model(0, L.MODEL_HARMONIC, [1.0, w0, 0.0])
    rng = np.random.default_rng(1)
    r0 = rng.normal(0, 0.3, size=(8, P)); v0 = rng.normal(0, 0.5, size=(8, P))
    E.load_initial(r0, v0, np.zeros((8, 2, 2)))
    E.run(0.0, tend, key)
    g = E.download(L.FIELD_R)
    E.close()
    U, w, _ = oracle.rp_normal_modes(P, P * T, dt)
    Om = np.sqrt(w ** 2 + w0 ** 2)
    nsteps = int(round(tend / dt))
    # the exact discrete map of one BCB step per normal mode: half kick by the bead potential,
    # Cayley step cay(dt A_k) = (I - dt A_k/2)^-1 (I + dt A_k/2) of the free ring polymer
    # (ring_polymer.jl:71-82), half kick
    K = np.array([[1.0, 0.0], [-w0 ** 2 * dt / 2, 1.0]])
    maps = []
    for wk in w:
        Ak = np.array([[0.0, 1.0], [-wk ** 2, 0.0]])
        Ck = np.linalg.solve(np.eye(2) - dt * Ak / 2, np.eye(2) + dt * Ak / 2)
        maps.append(np.linalg.matrix_power(K @ Ck @ K, nsteps))
    for i in range(8):
        rn0, vn0 = U.T @ r0[i], U.T @ v0[i]
        discrete = U @ np.array([Mk[0, 0] * a + Mk[0, 1] * b for Mk, a, b in zip(maps, rn0, vn0)])
        assert np.abs(g[i] - discrete).max() < 1e-10
        continuum = U @ (rn0 * np.cos(Om * tend) + vn0 / Om * np.sin(Om * tend))
        assert np.abs(g[i] - continuum).max() < 1e-3   # second-order integrator error at dt = 0.002


def test_rpmd_fp32_centroid_statistics(oracle, key):
    """FP32 beads: the mean centroid trajectory of 20000 thermal ring polymers (P = 32) matches
    the FP64 path within 1e-3 bohr."""
    from nqcd_mi355x import _lib as L
    ntraj, P, nsteps = 20000, 32, 2000
    out = []
    for prec in (0, 1):
        E = _ens(L, L.METHOD_RPMD, P, 1, precision=prec)
        E.sample_initial(L.SAMPLER_RING_POLYMER_THERMAL, ntraj, [-3.87, np.sqrt(0.1 / MASS), BETA], key)
        E.run(0.0, float(nsteps), key, save_every=100)
        out.append(E.reduce(L.OUT_CENTROID_POSITION, L.RED_MEAN, E.nsave))
        E.close()
    assert np.abs(out[0] - out[1]).max() < 1e-3


def _hartley(P):
    j = np.arange(P)
    a = 2 * np.pi * ((j[:, None] * j[None, :]) % P) / P
    return (np.cos(a) + np.sin(a)) / np.sqrt(P)


def test_harmonic_thermal_sampler_moments(key):
    """The harmonic-approximation sampler (include/nqcd_mi355x.h, SAMPLER_RING_POLYMER_THERMAL): normal mode k
    of the beads has mean sqrt(P) x0 delta_k0 and variance 1/(beta_P m (w_k^2 + w_well^2)), bead velocities
    variance 1/(beta_P m), beta_P = beta / P; 5 sigma of the moment estimates."""
    from nqcd_mi355x import _lib as L
    P, ntraj, x0, ww = 8, 50000, -3.87, np.sqrt(0.1 / MASS)
    r, v = _thermal_beads(L, P, ntraj, key, x0=x0)
    y = r @ _hartley(P).T
    bP = BETA / P
    wk = 2 * (P / BETA) * np.sin(np.arange(P) * np.pi / P)
    var = 1 / (bP * MASS * (wk ** 2 + ww ** 2))
    mean = np.r_[np.sqrt(P) * x0, np.zeros(P - 1)]
    assert np.all(np.abs(y.mean(axis=0) - mean) < 5 * np.sqrt(var / ntraj))
    assert np.all(np.abs(y.var(axis=0) / var - 1) < 5 * np.sqrt(2.0 / ntraj))
    assert abs(v.var() * bP * MASS - 1) < 5 * np.sqrt(2.0 / v.size) and abs(v.mean()) < 5 * np.sqrt(1 / (bP * MASS) / v.size)


def _pimc_device(L, P, ntraj, key, model, params, mass, T, nstates=1):
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_RPMD, 1, 1, 1, [mass], 1.0, algorithm=L.ALG_BCB, nbeads=P, temperature=T)
    E.set_model(0, model[0], model[1])
    E.sample_initial(L.SAMPLER_RING_POLYMER_PIMC, ntraj, params, key)
    E.run(0.0, 0.0, key)
    r, v = E.download(L.FIELD_R), E.download(L.FIELD_V)
    E.close()
    return r, v


@pytest.mark.parametrize("P,width,segment", [(1, 0, 1.0), (4, 0, 1.0), (8, 1, 0.5), (64, 0, 1.0)])
def test_pimc_parity(oracle, key, P, width, segment):
    """Device path-integral Monte Carlo chains (MetropolisHastings.jl:170-258, the reference's second test of
    the last proposal included) against the oracle's chains on the same Philox draws: final beads to 1e-9
    and the velocities, on the double well of the ring-polymer workload."""
    from nqcd_mi355x import _lib as L
    ntraj, passes, x0, delta = 300, 150, -3.87, 2.0
    print(f"PIMC parity P={P} width={width} segment={segment}", flush=True)
    r, v = _pimc_device(L, P, ntraj, key, (L.MODEL_DW_GAUSSIAN, DW1), [x0, delta, passes, segment, width], MASS,
                        1 / BETA)
    m = oracle.make_model(oracle.MODEL_DW_GAUSSIAN, 1, 1, DW1)
    xc, vc, acc = oracle.pimc(m, np.full((ntraj, P), x0), MASS, 1 / BETA, delta, passes, segment=segment, width=width,
                              key=key, reassess=1)
    assert acc.min() > 0
    assert np.allclose(r, xc, rtol=1e-9, atol=1e-9)
    assert np.allclose(v, vc, rtol=1e-12, atol=1e-15)


def test_pimc_path_integral_law_on_device(key):
    """Without the second test and with the free ring polymer's mode widths (params width = 1, reassess = 0)
    the chain samples the discretised path-integral distribution: on a harmonic well every normal mode has
    variance 1/(beta_P m (w^2 + w_k^2)) and the centroid mean 0 (5 sigma), P = 8, 20000 chains."""
    from nqcd_mi355x import _lib as L
    P, ntraj, mass, w, T = 8, 20000, 1.3, 0.9, 1.0 / 8
    r, v = _pimc_device(L, P, ntraj, key, (L.MODEL_HARMONIC, [mass, w, 0.0]), [0.0, 1.5, 300, 1.0, 1, 0], mass, T)
    y = r @ _hartley(P).T
    wk = 2 * P * T * np.sin(np.arange(P) * np.pi / P)
    var = 1 / (mass * (w ** 2 + wk ** 2) / (P * T))
    z = (y.var(axis=0) - var) / (var * np.sqrt(2.0 / ntraj))
    assert np.all(np.abs(z) < 5), z
    assert np.all(np.abs(y.mean(axis=0)) < 5 * np.sqrt(var / ntraj))
    assert abs(v.var() / (P * T / mass) - 1) < 5 * np.sqrt(2.0 / v.size)


@pytest.mark.parametrize("P", [4, 8, 16, 32])
def test_rp_layouts_agree_fp32(key, P, monkeypatch):
    """The two FP32 layouts (one ring polymer per lane; P lanes per ring polymer with the electronics on
    their own lanes) integrate the same ring polymers: RPSH over 1500 steps from PIMC beads ends in the same
    active state for at least 97% of 4096 ring polymers, and RPMD centroids agree to FP32 accuracy."""
    from nqcd_mi355x import _lib as L
    ntraj, nsteps = 4096, 1500
    out = {}
    for layout in ("bead", "lane"):
        monkeypatch.setenv("NQCD_MI355X_RP_LAYOUT", layout)
        for method, ns in ((L.METHOD_RPSH, 2), (L.METHOD_RPMD, 1)):
            E = _ens(L, method, P, ns, precision=1)
            E.sample_initial(L.SAMPLER_RING_POLYMER_PIMC, ntraj, [-3.87, 2.0, 100, 1.0, 1, 0], key, initial_state=0)
            E.run(0.0, float(nsteps if ns == 2 else 200), key, save_every=50)
            out[layout, ns] = (E.download(L.FIELD_STATE), E.download(L.FIELD_SAVE_CENTROID))
            E.close()
    assert np.mean(out["bead", 2][0] == out["lane", 2][0]) > 0.97
    cb, cl = out["bead", 1][1], out["lane", 1][1]
    assert np.allclose(cb, cl, rtol=0, atol=1e-4 * np.abs(cb).max()), np.abs(cb - cl).max()


def test_rpsh_fp32_beads_follow_fp64(key):
    """FP32 beads with the FP64 centroid electronics on their own lanes: the final active states of 4096
    thermal ring polymers (P = 8) agree with the FP64 kernel's on at least 97% of them."""
    from nqcd_mi355x import _lib as L
    P, ntraj, nsteps = 8, 4096, 1500
    states = []
    for prec in (0, 1):
        E = _ens(L, L.METHOD_RPSH, P, 2, precision=prec)
        E.sample_initial(L.SAMPLER_RING_POLYMER_PIMC, ntraj, [-3.87, 2.0, 100, 1.0, 1, 0], key, initial_state=0)
        E.run(0.0, float(nsteps), key)
        states.append(E.download(L.FIELD_STATE))
        E.close()
    assert np.mean(states[0] == states[1]) > 0.97
