"""Classical master equation surface hopping on the device (CME / BCME, cme.jl:26-182) against the FP64
oracle: the broadening integrals of the BCME force (device: integration by parts, then a closed form or panels;
oracle: adaptive Gauss-Kronrod to 1e-13), CME trajectories bit-for-bit in their hop sequences with positions
and energies to 1e-9, BCME trajectories to the accuracy of the device quadrature, and the reference's own
phonon-relaxation test (test/Dynamics/cme.jl:57-88) run through run_dynamics against its golden curve."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _ah(hw, Ed, g, Gam):
    # V11 = hw x^2 / 2, V22 = V11 + Ed + sqrt(2) g x, V12 = sqrt(Gam / 2 pi) (test/Dynamics/cme.jl:36-44)
    return [hw, Ed, np.sqrt(2) * g, np.sqrt(Gam / (2 * np.pi))]


@pytest.mark.parametrize("mu,kT,Gam,W", [(0.0, 0.01, 0.1, 100.0), (0.05, 0.03, 0.003, 10.0), (0.0, 0.001, 0.5, 5.0),
                                          (0.1, 0.2, 0.02, 100.0), (0.0, 1e-4, 1e-3, 1.0),
                                          (0.05, 0.03, 0.003, 1.0), (0.0, 0.01, 0.1, 0.5)])
def test_broadening_kernel_matches_oracle(oracle, mu, kT, Gam, W):
    """A band holding +-40 kT of mu takes the closed form (complex digamma: Stirling's series after shifting the
    argument to |z| >= 10, z = 1/2 + (Gamma/2 + i (mu - h)) / (2 pi kT), from near 1/2 to |z| in the hundreds
    over these cases); the last two cases, narrower bands, take eight-point panels no wider than Gamma / 2
    around h, whose error the nearest poles of the integrands (h +- i Gamma / 2) bound near 1e-10 of the
    integral.  The test allows 1e-9."""
    from nqcd_mi355x import _lib as L
    hs = np.concatenate([np.linspace(-1.0, 1.0, 81), [mu, mu + 1e-3, mu - 0.3 * kT, 0.5 * W, -0.5 * W]])
    print(f"broadening: mu={mu} kT={kT} Gamma={Gam} W={W}", flush=True)
    E = L.Ensemble(0)
    n, n2 = E.test_broadening(hs, mu, kT, Gam, W)
    E.close()
    rn = np.array([oracle.cme_broadening(h, mu, 1 / kT, Gam, W, 0) for h in hs])
    rn2 = np.array([oracle.cme_broadening(h, mu, 1 / kT, Gam, W, 1) for h in hs])
    assert np.allclose(n, rn, rtol=1e-9, atol=1e-12), np.abs(n - rn).max()
    assert np.allclose(n2, rn2, rtol=1e-9, atol=1e-12 * Gam), np.abs(n2 - rn2).max()


def _device_run(L, method, params, mass, kT, dt, nsteps, save_every, ntraj, prm, key, bandwidth=0.0, state0=0):
    E = L.Ensemble(0)
    E.set_problem(method, 2, 1, 1, [mass], dt, algorithm=L.ALG_TSIT5, rescaling=L.RESCALE_OFF, temperature=kT,
                  bandwidth=bandwidth)
    E.set_model(0, L.MODEL_ANDERSON_HOLSTEIN_1, params)
    E.sample_initial(L.SAMPLER_GAUSSIAN, ntraj, prm, key, initial_state=state0)
    E.run(0.0, dt * nsteps, key, save_every=save_every)
    g = {f: E.download(getattr(L, "FIELD_" + f)) for f in ("R", "V", "STATE", "NHOPS", "STEPS", "SAVE_R", "SAVE_EKIN",
                                                             "SAVE_EPOT", "SAVE_STATE")}
    g["red_ekin"] = E.reduce(L.OUT_KINETIC_ENERGY, L.RED_MEAN, E.nsave)
    g["red_x"] = E.reduce(L.OUT_POSITION, L.RED_MEAN, E.nsave)
    if method == L.METHOD_CME:
        g["red_etot"] = E.reduce(L.OUT_TOTAL_ENERGY, L.RED_SUM, E.nsave)
    E.close()
    return g


@pytest.mark.parametrize("state0", [0, 1])
def test_cme_trajectories_match_oracle(oracle, key, state0):
    from nqcd_mi355x import _lib as L
    hw, Gam, kT, g, Ed = 0.003, 0.003, 0.03, 0.005, 0.0
    mass, dt, nsteps, se, ntraj = 1 / hw, 5.0, 4000, 40, 512
    prm = [0.0, np.sqrt(2 * kT / hw), 0.0, np.sqrt(2 * kT / mass)]
    print(f"CME parity: state0={state0} ntraj={ntraj} steps={nsteps}", flush=True)
    dev = _device_run(L, L.METHOD_CME, _ah(hw, Ed, g, Gam), mass, kT, dt, nsteps, se, ntraj, prm, key, state0=state0)
    r0, v0 = oracle.sample_gaussian(ntraj, 1, *prm, key)
    m = oracle.make_model(oracle.MODEL_ANDERSON_HOLSTEIN_1, 2, 1, _ah(hw, Ed, g, Gam))
    c = oracle.cme_run(0, m, mass, kT, r0[:, 0], v0[:, 0], np.full(ntraj, state0, np.int32), dt=dt, nsteps=nsteps,
                       save_every=se, key=key)
    assert c["nhops"].sum() > ntraj
    assert np.array_equal(dev["STATE"], c["state"])
    assert np.array_equal(dev["NHOPS"], c["nhops"])
    assert np.array_equal(dev["SAVE_STATE"], c["save_state"])
    assert np.all(dev["STEPS"] == nsteps)
    assert np.allclose(dev["R"][:, 0], c["x"], rtol=1e-9, atol=1e-9)
    assert np.allclose(dev["V"][:, 0], c["v"], rtol=1e-9, atol=1e-12)
    assert np.allclose(dev["SAVE_R"].reshape(ntraj, -1), c["save_x"], rtol=1e-9, atol=1e-9)
    assert np.allclose(dev["SAVE_EKIN"], c["ekin"], rtol=1e-9, atol=1e-14)
    assert np.allclose(dev["SAVE_EPOT"], c["epot"], rtol=1e-9, atol=1e-14)
    assert np.allclose(dev["red_ekin"], c["ekin"].mean(axis=0), rtol=1e-12, atol=1e-16)
    assert np.allclose(dev["red_x"], c["save_x"].mean(axis=0), rtol=1e-9, atol=1e-9)
    assert np.allclose(dev["red_etot"], (c["ekin"] + c["epot"]).sum(axis=0), rtol=1e-9, atol=1e-12)


def test_bcme_trajectories_match_oracle(oracle, key):
    """BCME force differs from the oracle's by the quadrature error (relative 1e-10 or below), so positions
    agree to 1e-6 over 3000 steps and the hop sequences agree except where a uniform falls within that
    error of the hop probability."""
    from nqcd_mi355x import _lib as L
    hw, Gam, kT, g = 0.003, 0.1, 0.01, 0.015
    Ed = g ** 2 / hw
    mass, dt, nsteps, se, ntraj, W = 1 / hw, 2.0, 3000, 50, 256, 100.0
    x1 = -np.sqrt(2) * g / hw
    prm = [x1, np.sqrt(kT / hw), 0.0, np.sqrt(kT / mass)]
    print(f"BCME parity: ntraj={ntraj} steps={nsteps}", flush=True)
    dev = _device_run(L, L.METHOD_BCME, _ah(hw, Ed, g, Gam), mass, kT, dt, nsteps, se, ntraj, prm, key, bandwidth=W)
    r0, v0 = oracle.sample_gaussian(ntraj, 1, *prm, key)
    m = oracle.make_model(oracle.MODEL_ANDERSON_HOLSTEIN_1, 2, 1, _ah(hw, Ed, g, Gam))
    c = oracle.cme_run(1, m, mass, kT, r0[:, 0], v0[:, 0], np.zeros(ntraj, np.int32), dt=dt, nsteps=nsteps,
                       bandwidth=W, save_every=se, key=key)
    same = np.all(dev["SAVE_STATE"] == c["save_state"], axis=1) & (dev["NHOPS"] == c["nhops"])
    assert c["nhops"].sum() > ntraj
    assert same.mean() >= 0.98, same.mean()
    assert np.allclose(dev["SAVE_R"].reshape(ntraj, -1)[same], c["save_x"][same], rtol=0, atol=1e-6)
    assert np.allclose(dev["V"][same, 0], c["v"][same], rtol=0, atol=1e-8)


def _golden(name):
    d = np.loadtxt(os.path.join(HERE, "golden", name), delimiter=",")
    o = np.argsort(d[:, 0])
    x, y = d[o, 0], d[o, 1]

    def itp(t):  # linear interpolation, linear extrapolation (Interpolations' Line())
        r = np.interp(t, x, y)
        lo, hi = t < x[0], t > x[-1]
        r[lo] = y[0] + (t[lo] - x[0]) * (y[1] - y[0]) / (x[1] - x[0])
        r[hi] = y[-1] + (t[hi] - x[-1]) * (y[-1] - y[-2]) / (x[-1] - x[-2])
        return r
    return itp


@pytest.mark.parametrize("T", [0.2, 5.0])
def test_cme_phonon_relaxation_run_dynamics(T):
    """test/Dynamics/cme.jl:57-88 through run_dynamics: 1000 trajectories from a thermal oscillator at
    kT_i = T kT on diabat 1, saveat 2/Gamma to 200/Gamma; mean kinetic energy / kT against the golden curve
    within 5 standard errors or 20% (the reference's tolerances)."""
    import nqcd_mi355x as N
    hw, Gam, kT, g, Ed = 0.003, 0.003, 0.03, 0.005, 0.0
    mass = 1 / hw
    sim = N.Simulation[N.CME](N.Atoms([mass]), N.AndersonHolsteinOneMode(hw, Ed, g, Gam), temperature=kT)
    kTi = kT * T
    r = N.Normal(0.0, np.sqrt(kTi / hw))
    dist = N.DynamicalDistribution(N.VelocityBoltzmann(kTi, mass), r, (1, 1)) * N.PureState(1, N.Diabatic())
    out = N.run_dynamics(sim, (0.0, 200 / Gam), dist, trajectories=1000, dt=2 / Gam / 133, saveat=2 / Gam,
                         output=(N.OutputKineticEnergy, N.OutputTotalEnergy, N.OutputDiscreteState),
                         ensemble_algorithm=N.EnsembleMI355X(seed=7 + int(10 * T)))
    ek = np.array([o["OutputKineticEnergy"] for o in out])
    avg = ek.mean(axis=0) / kT
    err = ek.std(axis=0, ddof=1) / np.sqrt(len(out)) / kT
    t = np.asarray(out[0]["Time"]) * Gam
    ref = _golden(f"cme_{T}T.csv")(t)
    ok = np.abs(ref - avg) <= np.maximum(5 * err, 0.2 * np.abs(ref))
    assert ok.all(), (t[~ok], avg[~ok], ref[~ok])
    st = np.array([o["OutputDiscreteState"] for o in out])
    assert set(np.unique(st)) <= {1, 2} and np.all(st[:, 0] == 1)


def test_bcme_position_relaxation_run_dynamics():
    """test/Dynamics/bcme.jl:57-95 through run_dynamics with MeanReduction on the device: mean x / x1 every
    100 a.u. to 10000 against the golden curve within 10 standard errors (from a per-trajectory run) or 10%."""
    import nqcd_mi355x as N
    hw, Gam, kT, g = 0.003, 0.1, 0.01, 0.015
    Ed = g ** 2 / hw
    mass = 1 / hw
    x1 = -np.sqrt(2) * g / hw
    sim = N.Simulation[N.BCME](N.Atoms([mass]), N.AndersonHolsteinOneMode(hw, Ed, g, Gam), temperature=kT,
                               bandwidth=100.0)
    dist = N.DynamicalDistribution(N.VelocityBoltzmann(kT, mass), N.Normal(x1, np.sqrt(kT / hw)), (1, 1)) \
        * N.PureState(1, N.Diabatic())
    alg = N.EnsembleMI355X(seed=11)
    per = N.run_dynamics(sim, (0.0, 10000.0), dist, trajectories=100, dt=2.0, saveat=100.0, output=N.OutputPosition,
                         ensemble_algorithm=alg)
    xs = np.array([[float(np.ravel(x)[0]) for x in o["OutputPosition"]] for o in per]) / x1
    mean = N.run_dynamics(sim, (0.0, 10000.0), dist, trajectories=100, dt=2.0, saveat=100.0, output=N.OutputPosition,
                          reduction=N.MeanReduction(), ensemble_algorithm=alg)
    dm = np.array([float(np.ravel(x)[0]) for x in mean["OutputPosition"]]) / x1
    assert np.allclose(dm, xs.mean(axis=0), rtol=1e-12, atol=1e-12)
    err = xs.std(axis=0, ddof=1) / np.sqrt(xs.shape[0])
    t = np.asarray(mean["Time"])
    ref = _golden("bcme.csv")(t)
    ok = np.abs(ref - dm) <= np.maximum(10 * err, 0.1 * np.abs(ref))
    assert ok.all(), (t[~ok], dm[~ok], ref[~ok])
