"""Pins the oracle's classical master equation surface hopping (CME, BCME; cme.jl:26-182) against the
reference's own golden data: phonon relaxation of the Anderson-Holstein impurity at three initial
temperatures (test/Dynamics/cme.jl:57-88, reference_data/{0.2T,2.0T,5.0T}.csv) and the BCME position
relaxation (test/Dynamics/bcme.jl:57-95, reference_data/bcme.csv), with the reference's tolerances; the
broadening integrals against scipy's adaptive quadrature; and the hop law on a frozen configuration."""
import os

import numpy as np
import pytest
from scipy.integrate import quad

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden(name):
    d = np.loadtxt(os.path.join(HERE, "golden", name), delimiter=",")
    o = np.argsort(d[:, 0])
    x, y = d[o, 0], d[o, 1]

    def itp(t):  # linear interpolation with linear extrapolation (Interpolations' Line())
        r = np.interp(t, x, y)
        lo, hi = t < x[0], t > x[-1]
        r[lo] = y[0] + (t[lo] - x[0]) * (y[1] - y[0]) / (x[1] - x[0])
        r[hi] = y[-1] + (t[hi] - x[-1]) * (y[-1] - y[-2]) / (x[-1] - x[-2])
        return r
    return itp


def _ah_model(oracle, hw, Ed, g, Gam):
    # V11 = hw x^2 / 2, V22 = V11 + Ed + sqrt(2) g x, V12 = sqrt(Gam / 2 pi) (test/Dynamics/cme.jl:36-44)
    return oracle.make_model(oracle.MODEL_ANDERSON_HOLSTEIN_1, 2, 1, [hw, Ed, np.sqrt(2) * g, np.sqrt(Gam / (2 * np.pi))])


@pytest.mark.parametrize("T", [0.2, 2.0, 5.0])
def test_cme_phonon_relaxation_golden(oracle, key, T):
    """test/Dynamics/cme.jl:57-88: 1000 trajectories from a thermal oscillator at kT_i = T kT on diabat 1,
    mean kinetic energy / kT every 2/Gamma to 200/Gamma against the golden curve within 5 standard errors
    or 20%."""
    hw, Gam, kT, g, Ed = 0.003, 0.003, 0.03, 0.005, 0.0
    mass = 1 / hw
    m = _ah_model(oracle, hw, Ed, g, Gam)
    ntraj, dt = 1000, 5.0
    kTi = kT * T
    sig_r = np.sqrt(kTi / hw)          # sqrt(1 / beta omega) in the reference's reduced units
    sig_v = np.sqrt(kTi / mass)        # VelocityBoltzmann(kTi, m)
    r0, v0 = oracle.sample_gaussian(ntraj, 1, 0.0, sig_r, 0.0, sig_v, key)
    tend, every = 200 / Gam, 2 / Gam
    nsteps = int(round(tend / dt))
    se = int(round(every / dt))
    out = oracle.cme_run(0, m, mass, kT, r0[:, 0], v0[:, 0], np.zeros(ntraj, np.int32), dt=dt, nsteps=nsteps,
                         save_every=se, key=key)
    avg = out["ekin"].mean(axis=0) / kT
    err = out["ekin"].std(axis=0, ddof=1) / np.sqrt(ntraj) / kT
    t = np.arange(out["ekin"].shape[1]) * se * dt * Gam
    ref = _golden(f"cme_{T}T.csv")(t)
    ok = np.abs(ref - avg) <= np.maximum(5 * err, 0.2 * np.abs(ref))
    assert ok.all(), (t[~ok], avg[~ok], ref[~ok])
    assert out["nhops"].sum() > ntraj


def test_bcme_position_relaxation_golden(oracle, key):
    """test/Dynamics/bcme.jl:57-95: 100 BCME trajectories (bandwidth 100) from the thermal distribution
    around the displaced minimum x1 = -sqrt(2) g / hw, mean x / x1 every 100 a.u. to 10000 against the
    golden curve within 10 standard errors or 10%."""
    hw, Gam, kT, g = 0.003, 0.1, 0.01, 0.015
    Ed = g ** 2 / hw
    mass = 1 / hw
    x1 = -np.sqrt(2) * g / hw
    m = _ah_model(oracle, hw, Ed, g, Gam)
    ntraj, dt = 100, 2.0
    r0, v0 = oracle.sample_gaussian(ntraj, 1, x1, np.sqrt(kT / hw), 0.0, np.sqrt(kT / mass), key)
    out = oracle.cme_run(1, m, mass, kT, r0[:, 0], v0[:, 0], np.zeros(ntraj, np.int32), dt=dt, nsteps=5000,
                         bandwidth=100.0, save_every=50, key=key)
    avg = (out["save_x"] / x1).mean(axis=0)
    err = (out["save_x"] / x1).std(axis=0, ddof=1) / np.sqrt(ntraj)
    t = np.arange(out["save_x"].shape[1]) * 50 * dt
    ref = _golden("bcme.csv")(t)
    ok = np.abs(ref - avg) <= np.maximum(10 * err, 0.1 * np.abs(ref))
    assert ok.all(), (t[~ok], avg[~ok], ref[~ok])


@pytest.mark.parametrize("h", [-0.3, 0.0, 0.04, 0.7])
def test_broadening_integrals_against_scipy(oracle, h):
    """evaluate_broadening / evaluate_force_broadening (cme.jl:167-182) against scipy's adaptive quadrature."""
    mu, beta, Gam, W = 0.0, 100.0, 0.1, 100.0
    A = lambda e: Gam / 2 / np.pi / ((e - h) ** 2 + (Gam / 2) ** 2)
    fe = lambda e: 1 / (1 + np.exp(np.clip(beta * (e - mu), -700, 700)))
    cuts = sorted({-W / 2, W / 2, h, mu})
    for force, k in ((0, lambda e: A(e) * fe(e)), (1, lambda e: (e - h) * A(e) * fe(e))):
        ref = sum(quad(k, a, b, epsabs=0, epsrel=1e-13, limit=500)[0] for a, b in zip(cuts[:-1], cuts[1:]))
        assert np.isclose(oracle.cme_broadening(h, mu, beta, Gam, W, force), ref, rtol=1e-11, atol=1e-15)


def test_cme_hop_law(oracle, key):
    """On a frozen configuration (no force: k = 0, c = 0) the state is a two-state Markov chain with hop
    probabilities Gamma f dt and Gamma (1 - f) dt per step (cme.jl:26-53): after many steps the fraction in
    state 2 is f (detailed balance), and the first-step hop fraction from state 1 is Gamma f dt."""
    Gam, kT, Ed, dt = 0.02, 0.01, 0.005, 1.0
    m = oracle.make_model(oracle.MODEL_ANDERSON_HOLSTEIN_1, 2, 1, [0.0, Ed, 0.0, np.sqrt(Gam / (2 * np.pi))])
    f = 1 / (1 + np.exp(Ed / kT))
    ntraj = 40000
    out = oracle.cme_run(0, m, 1.0, kT, np.zeros(ntraj), np.zeros(ntraj), np.zeros(ntraj, np.int32), dt=dt, nsteps=1,
                         key=key)
    p1 = np.mean(out["state"] == 1)
    assert abs(p1 - Gam * f * dt) < 5 * np.sqrt(Gam * f * dt / ntraj)
    out = oracle.cme_run(0, m, 1.0, kT, np.zeros(ntraj), np.zeros(ntraj), np.zeros(ntraj, np.int32), dt=dt, nsteps=2000,
                         key=key)
    pinf = np.mean(out["state"] == 1)
    assert abs(pinf - f) < 5 * np.sqrt(f * (1 - f) / ntraj), (pinf, f)
