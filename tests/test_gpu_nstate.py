"""FSSH and Ehrenfest with more than two electronic states (fssh.jl:96-121, electronic_dynamics.jl:104-130,
BABwithTsit5) against the FP64 oracle's general-n path: n-state Morse chains at n = 3, 4 and 8 over short
windows (identical hop sequences, positions and velocities to 1e-9, densities to 1e-9 plus the oracle's own
response to a 1e-13 relative change of the initial positions, since near-degenerate crossings amplify
rounding in the coherences), and one persistent launch mixing two-state Tully and three-state Morse
trajectories."""
import numpy as np
import pytest

import nstate_models as NS

pytestmark = pytest.mark.gpu


def _ens(L, method, n, precision=0):
    E = L.Ensemble(0)
    E.set_problem(method, n, 1, 1, [NS.MASS], 1.0, precision=precision)
    return E


def _initial(n, ntraj, start, seed):
    rng = np.random.default_rng(seed)
    r0 = rng.normal(3.0, 0.15, size=(ntraj, 1))
    v0 = np.sqrt(2 * rng.uniform(0.02, 0.04, size=(ntraj, 1)) / NS.MASS)
    sig0 = np.zeros((ntraj, n, n), complex)
    sig0[:, start, start] = 1.0
    return r0, v0, sig0


@pytest.mark.parametrize("n,start,nsteps", [(3, 0, 1200), (4, 1, 1000), (8, 3, 400)])
@pytest.mark.parametrize("method", [0, 1])
def test_nstate_fp64_parity(oracle, key, n, start, nsteps, method):
    from nqcd_mi355x import _lib as L
    table = NS.three_state() if n == 3 else NS.morse_table(n)
    ntraj = 193
    r0, v0, sig0 = _initial(n, ntraj, start, n)
    st0 = np.full(ntraj, start, np.int32)
    print(f"n-state parity: n={n} method={method} ntraj={ntraj} steps={nsteps}", flush=True)
    E = _ens(L, method, n)
    E.set_model(0, L.MODEL_MORSE_N, [n] + table)
    E.load_initial(r0, v0, sig0, state=st0)
    E.run(0.0, float(nsteps), key, save_every=100)
    g = {f: E.download(getattr(L, "FIELD_" + f)) for f in ("R", "V", "SIGMA_RE", "SIGMA_IM", "STATE", "NHOPS",
                                                             "SAVE_POP_AD", "SAVE_POP_DI", "SAVE_STATE")}
    E.close()
    m = oracle.make_model(oracle.MODEL_MORSE_N, n, 1, table)
    run = lambda r: oracle.run_ensemble(method, m, NS.MASS, r, v0, sig0, state0=st0, dt=1.0, nsteps=nsteps, key=key,
                                        save_every=100, nsave=nsteps // 100 + 1)
    c, cp = run(r0), run(r0 * (1 + 1e-13))
    if method == 0:
        assert c["nhops"].sum() > 0
        assert np.array_equal(g["STATE"], c["state"])
        assert np.array_equal(g["NHOPS"], c["nhops"])
        assert np.array_equal(g["SAVE_STATE"], c["save_state"])
    assert np.allclose(g["R"], c["r"], rtol=1e-9, atol=1e-10)
    assert np.allclose(g["V"], c["v"], rtol=1e-9, atol=1e-13)
    for f, k in (("SIGMA_RE", "sig_re"), ("SIGMA_IM", "sig_im"), ("SAVE_POP_AD", "pop_ad"), ("SAVE_POP_DI", "pop_di")):
        sens = np.abs(cp[k] - c[k]).reshape(ntraj, -1).max(axis=1)
        tol = 1e-9 + sens
        err = np.abs(g[f] - c[k]).reshape(ntraj, -1).max(axis=1)
        bad = np.nonzero(err > tol)[0]
        assert bad.size == 0, (f, bad[:5], err[bad[:5]], sens[bad[:5]])


@pytest.mark.parametrize("method", [0, 1])
def test_mixed_state_counts_one_launch(oracle, key, method):
    """Slot 0 Tully 1 (two states), slot 1 three-state Morse, interleaved in one ensemble and one launch of
    the three-state kernel: each trajectory matches the oracle run of its own model and state count, and
    the two-state trajectories keep zero population in the third state."""
    from nqcd_mi355x import _lib as L
    ntraj, nsteps = 300, 800
    tully = [0.01, 1.6, 0.005, 1.0]
    mid = (np.arange(ntraj) % 3 == 1).astype(np.int32)  # a third of the trajectories on the Morse model
    rng = np.random.default_rng(4)
    r0 = np.where(mid[:, None] == 1, rng.normal(3.0, 0.15, (ntraj, 1)), rng.normal(-4.0, 0.5, (ntraj, 1)))
    v0 = np.where(mid[:, None] == 1, np.sqrt(2 * 0.03 / NS.MASS), rng.uniform(10, 30, (ntraj, 1)) / NS.MASS)
    sig0 = np.zeros((ntraj, 3, 3), complex); sig0[:, 0, 0] = 1.0
    st0 = np.zeros(ntraj, np.int32)
    print(f"mixed launch: method={method}", flush=True)
    E = _ens(L, method, 3)
    E.set_model(0, L.MODEL_TULLY1, tully)
    E.set_model(1, L.MODEL_MORSE_N, [3] + NS.three_state())
    E.load_initial(r0, v0, sig0, state=st0, model_id=mid)
    st = E.run(0.0, float(nsteps), key, save_every=100)
    assert st["launches"] <= 2
    g = {f: E.download(getattr(L, "FIELD_" + f)) for f in ("R", "V", "SIGMA_RE", "STATE", "SAVE_POP_DI")}
    E.close()
    for k, (model, nst) in enumerate(((oracle.make_model(0, 2, 1, tully), 2),
                                      (oracle.make_model(oracle.MODEL_MORSE_N, 3, 1, NS.three_state()), 3))):
        sel = np.nonzero(mid == k)[0]
        c = oracle.run_ensemble(method, model, NS.MASS, r0[sel], v0[sel], sig0[sel][:, :nst, :nst], state0=st0[sel],
                                dt=1.0, nsteps=nsteps, key=key, save_every=100, nsave=nsteps // 100 + 1, traj_ids=sel)
        assert np.allclose(g["R"][sel], c["r"], rtol=1e-9, atol=1e-10)
        assert np.allclose(g["V"][sel], c["v"], rtol=1e-9, atol=1e-13)
        assert np.allclose(g["SIGMA_RE"][sel][:, :nst, :nst], c["sig_re"], rtol=1e-9, atol=1e-10)
        assert np.allclose(g["SAVE_POP_DI"][sel][:, :, :nst], c["pop_di"], rtol=1e-9, atol=1e-10)
        if method == 0:
            assert np.array_equal(g["STATE"][sel], c["state"])
        if nst == 2:
            assert np.all(g["SIGMA_RE"][sel][:, 2, :] == 0) and np.all(g["SAVE_POP_DI"][sel][:, :, 2] == 0)


def test_nstate_diabatic_start_fp64(oracle, key):
    """Device-sampled Gaussian initial conditions with a diabatic initial state: sigma = U^T e_d e_d^T U at r0
    and the active state drawn from its diagonal (fssh.jl:54) in the kernel, against the oracle run from the
    oracle's own diabatic_to_adiabatic on the same draws: identical active states and hop sequences."""
    from nqcd_mi355x import _lib as L
    ntraj, nsteps, n = 1024, 1500, 3
    prm = [3.0, 0.15, np.sqrt(2 * 0.03 / NS.MASS), 0.0]
    print("n-state diabatic start, FP64", flush=True)
    E = _ens(L, 0, n)
    E.set_model(0, L.MODEL_MORSE_N, [n] + NS.three_state())
    E.sample_initial(L.SAMPLER_GAUSSIAN, ntraj, prm, key, initial_state=1, state_basis=1)
    E.run(0.0, float(nsteps), key, save_every=500)
    g = {f: E.download(getattr(L, "FIELD_" + f)) for f in ("R", "V", "STATE", "NHOPS", "SAVE_STATE", "SAVE_POP_DI")}
    E.close()
    m = oracle.make_model(oracle.MODEL_MORSE_N, n, 1, NS.three_state())
    r0, v0 = oracle.sample_gaussian(ntraj, 1, prm[0], prm[1], prm[2], prm[3], key)
    sig0 = np.stack([oracle.diabatic_to_adiabatic(m, r0[i:i + 1], 1)[0] for i in range(ntraj)])
    c = oracle.run_ensemble(0, m, NS.MASS, r0, v0, sig0, state0=np.full(ntraj, -1, np.int32), dt=1.0, nsteps=nsteps,
                            key=key, save_every=500, nsave=4)   # state -1: drawn from diag(sigma)
    assert np.array_equal(g["SAVE_STATE"][:, 0], c["save_state"][:, 0])   # the drawn initial states
    assert np.array_equal(g["STATE"], c["state"])
    assert np.array_equal(g["NHOPS"], c["nhops"])
    assert np.allclose(g["R"], c["r"], rtol=1e-9, atol=1e-10)
    assert np.allclose(g["V"], c["v"], rtol=1e-9, atol=1e-13)
    assert np.allclose(g["SAVE_POP_DI"][:, 0], c["pop_di"][:, 0], rtol=1e-9, atol=1e-10)


def test_nstate_fp32_nuclei_statistics(oracle, key):
    """The FP32-nuclei kernel (FP64 electronics) from uploaded adiabatic initial conditions: final active-state
    populations of a three-state ensemble within 5 sigma of the FP64 oracle on the same initial conditions."""
    from nqcd_mi355x import _lib as L
    ntraj, nsteps, n = 4096, 1500, 3
    r0, v0, sig0 = _initial(n, ntraj, 1, 11)
    st0 = np.full(ntraj, 1, np.int32)
    print("n-state FP32 nuclei statistics", flush=True)
    E = _ens(L, 0, n, precision=L.PREC_FP32_NUCLEI)
    E.set_model(0, L.MODEL_MORSE_N, [n] + NS.three_state())
    E.load_initial(r0, v0, sig0, state=st0)
    E.run(0.0, float(nsteps), key)
    gs = E.download(L.FIELD_STATE)
    E.close()
    m = oracle.make_model(oracle.MODEL_MORSE_N, n, 1, NS.three_state())
    c = oracle.run_ensemble(0, m, NS.MASS, r0, v0, sig0, state0=st0, dt=1.0, nsteps=nsteps, key=key)
    pg = np.bincount(gs, minlength=n) / ntraj
    pc = np.bincount(c["state"], minlength=n) / ntraj
    sig = np.sqrt(pc * (1 - pc) / ntraj) + 1e-3
    assert np.all(np.abs(pg - pc) < 5 * sig), (pg, pc)
