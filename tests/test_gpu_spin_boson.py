"""GPU spin-boson (100-mode bath) FSSH / Ehrenfest against the FP64 oracle and the
reference's golden curve (test/Dynamics/ehrenfest.jl:110-141)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
N = 100


def _debye(oracle):
    wc, lam = 5.0, 0.625  # eta = 0.25: lambda = eta wc / 2
    j = np.arange(1, N + 1)
    w = wc * np.tan(np.pi / 2 * (1 - j / (N + 1)))
    c = np.sqrt(2 * lam / (N + 1)) * w
    return w, c


def _wigner(w, beta):
    t = np.tanh(beta * w / 2)
    return np.sqrt(1 / (2 * w * t)), np.sqrt(w / (2 * t))


def _gpu(L, method, eps, Delta, w, c, sdr, sdv, ntraj, nsteps, dt, key, precision=0, save_every=1):
    E = L.Ensemble(0)
    E.set_problem(method, 2, 1, N, np.ones(N), dt, precision=precision)
    E.set_model(0, L.MODEL_SPIN_BOSON, np.concatenate([[eps, Delta], w, c]))
    prm = np.concatenate([np.zeros(N), sdr, np.zeros(N), sdv])
    E.sample_initial(L.SAMPLER_HARMONIC_WIGNER, ntraj, prm, key, initial_state=0, state_basis=1)
    E.run(0.0, nsteps * dt, key, save_every=save_every)
    out = {"r": E.download(L.FIELD_R), "v": E.download(L.FIELD_V), "state": E.download(L.FIELD_STATE),
           "sig": E.download(L.FIELD_SIGMA_RE) + 1j * E.download(L.FIELD_SIGMA_IM),
           "save_state": E.download(L.FIELD_SAVE_STATE), "pop_di": E.download(L.FIELD_SAVE_POP_DI),
           "nhops": E.download(L.FIELD_NHOPS)}
    corr = E.reduce(L.OUT_POPULATION_CORRELATION, L.RED_MEAN, 4 * E.nsave).reshape(E.nsave, 4)
    out["sz"] = corr[:, 0] - corr[:, 2]  # P1(0)P1(t) - P1(0)P2(t), column-major 2x2
    E.close()
    return out


def _cpu(oracle, method, eps, Delta, w, c, sdr, sdv, ntraj, nsteps, dt, key, save_every=1):
    m = oracle.make_model(oracle.MODEL_SPIN_BOSON, 2, N, [eps, Delta], omega=w, cbath=c)
    r0, v0 = oracle.sample_gaussian(ntraj, N, 0.0, sdr[None, :], 0.0, sdv[None, :], key)
    sig0 = oracle.diabatic_to_adiabatic(m, r0, 0)
    nsave = nsteps // save_every + 1
    out = oracle.run_ensemble(method, m, 1.0, r0, v0, sig0, state0=None if method == 1 else np.full(ntraj, -1, np.int32),
                              dt=dt, nsteps=nsteps, save_every=save_every, nsave=nsave, key=key)
    P = out["pop_di"]
    out["sz"] = np.mean(P[:, 0:1, 0] * P[:, :, 0] - P[:, 0:1, 0] * P[:, :, 1], axis=0)
    return out


@pytest.mark.parametrize("method", [1, 0])
def test_spin_boson_fp64_window_parity(oracle, key, method):
    from nqcd_mi355x import _lib as L
    w, c = _debye(oracle)
    sdr, sdv = _wigner(w, 5.0)
    ntraj, nsteps, dt = 256, 300, 0.002   # omega_max = 321: velocity Verlet needs omega_max dt < 2
    g = _gpu(L, method, 1.0, 1.0, w, c, sdr, sdv, ntraj, nsteps, dt, key)
    cc = _cpu(oracle, method, 1.0, 1.0, w, c, sdr, sdv, ntraj, nsteps, dt, key)
    if method == 0:
        assert cc["nhops"].sum() > 10
        assert np.array_equal(g["save_state"], cc["save_state"])
    assert np.allclose(g["r"], cc["r"], rtol=1e-9, atol=1e-11)
    assert np.allclose(g["v"], cc["v"], rtol=1e-9, atol=1e-11)
    assert np.allclose(g["sig"], cc["sig_re"] + 1j * cc["sig_im"], rtol=1e-9, atol=1e-11)
    assert np.allclose(g["pop_di"], cc["pop_di"], rtol=1e-9, atol=1e-11)
    assert np.allclose(g["sz"], cc["sz"], atol=1e-11)


def test_debye_workload_fp32_matches_fp64_oracle(oracle, key):
    """Ehrenfest, Debye bath (eps = Delta = 1, wc = 5, beta = 5, 100 modes) to t = 20: <sigma_z(t)> of
    the FP32-bath path within 0.01 of the FP64 oracle on the same 1000 trajectories.  The highest
    discretised mode is omega_1 = wc tan(pi/2 * 100/101) = 321, so the step is dt = 0.002
    (omega_1 dt = 0.64); dt = 0.01 puts omega_1 dt = 3.2 beyond velocity Verlet's stability limit of 2."""
    from nqcd_mi355x import _lib as L
    w, c = _debye(oracle)
    sdr, sdv = _wigner(w, 5.0)
    ntraj, nsteps, dt = 1000, 10000, 0.002
    g = _gpu(L, 1, 1.0, 1.0, w, c, sdr, sdv, ntraj, nsteps, dt, key, precision=1, save_every=50)
    cc = _cpu(oracle, 1, 1.0, 1.0, w, c, sdr, sdv, ntraj, nsteps, dt, key, save_every=50)
    # per-trajectory samples of P1(0) [P1(t) - P2(t)]
    sg = g["pop_di"][:, 0:1, 0] * (g["pop_di"][:, :, 0] - g["pop_di"][:, :, 1])
    sc = cc["pop_di"][:, 0:1, 0] * (cc["pop_di"][:, :, 0] - cc["pop_di"][:, :, 1])
    assert np.allclose(sg.mean(axis=0), g["sz"], atol=1e-12)
    diff = np.abs(g["sz"] - cc["sz"])
    t = np.arange(len(diff)) * 50 * dt
    print("max |FP32 - FP64| <sigma_z>: t <= 2:", diff[t <= 2].max(), " all t:", diff.max())
    # while the FP32 and FP64 copies of each trajectory stay together the curves agree closely
    assert diff[t <= 2].max() < 0.005
    # later the (chaotic) trajectories decorrelate; the ensemble means then agree within the
    # statistical error of the difference of two 1000-trajectory means
    sigma = np.sqrt((sg.var(axis=0) + sc.var(axis=0)) / ntraj)
    assert np.all(diff <= 4 * sigma + 1e-12), (diff / np.maximum(sigma, 1e-300)).max()


def test_gao_saller_golden_curve_through_run_dynamics(key):
    """The reference's spin-boson Ehrenfest test (Ohmic wc = 2.5, alpha = 0.09, eps = 0, Delta = 1,
    beta = 5, 100 modes, dt = 0.1, 500 trajectories, saveat 0:0.1:20) through the mirrored
    run_dynamics API on the FP32-bath path: atol 0.2 against gao_saller_jctc_2020_fig2b.csv."""
    import nqcd_mi355x as NQ
    model = NQ.SpinBoson(NQ.OhmicSpectralDensity(2.5, 0.09), N, 0.0, 1.0)
    atoms = NQ.Atoms([1.0] * N)
    beta = 5.0
    position = np.array([[NQ.PositionHarmonicWigner(w, beta, 1) for w in model.omega]], dtype=object)
    velocity = np.array([[NQ.VelocityHarmonicWigner(w, beta, 1) for w in model.omega]], dtype=object)
    dist = NQ.DynamicalDistribution(velocity, position, (1, N)) * NQ.PureState(1)
    sim = NQ.Simulation[NQ.Ehrenfest](atoms, model, precision="fp32_nuclei")
    out = NQ.PopulationCorrelationFunction(sim, NQ.Diabatic())
    res = NQ.run_dynamics(sim, (0.0, 20.0), dist, saveat=0.1, trajectories=500, output=out,
                          reduction=NQ.MeanReduction(), dt=0.1)
    result = np.array([p[0, 0] - p[0, 1] for p in res["PopulationCorrelationFunction"]])
    data = np.loadtxt(os.path.join(HERE, "golden", "gao_saller_jctc_2020_fig2b.csv"), delimiter=",")
    o = np.argsort(data[:, 0])
    tg, yg = data[o, 0], data[o, 1]
    times = np.asarray(res["Time"])
    ref = np.interp(times, tg, yg)
    lo, hi = times < tg[0], times > tg[-1]
    ref[lo] = yg[0] + (times[lo] - tg[0]) * (yg[1] - yg[0]) / (tg[1] - tg[0])
    ref[hi] = yg[-1] + (times[hi] - tg[-1]) * (yg[-1] - yg[-2]) / (tg[-1] - tg[-2])
    assert len(result) == 201
    assert np.all(np.abs(result - ref) < 0.2), np.abs(result - ref).max()
