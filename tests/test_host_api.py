"""Host-side mirror of the reference API (no device calls): model discretisations,
distribution moments and argument validation of run_dynamics."""
import numpy as np
import pytest

import nqcd_mi355x as N


def test_spectral_density_discretisations_match_docs():
    """docs/src/NQCModels/systembathmodels.md:54-66 (Ohmic) and :89-96 (Debye)."""
    w, c = N.OhmicSpectralDensity(2.5, 0.09).discretise(100)
    assert np.isclose(w[0], -2.5 * np.log(1 - 1 / 101)) and np.isclose(w[-1], -2.5 * np.log(1 / 101))
    assert np.allclose(c, np.sqrt(0.09 * 2.5 / 101) * w)
    w, c = N.DebyeSpectralDensity(5.0, 0.625).discretise(100)
    assert np.isclose(w[49], 5.0 * np.tan(np.pi / 2 * (1 - 50 / 101)))
    assert np.allclose(c, np.sqrt(2 * 0.625 / 101) * w)
    # Debye discretisation reproduces the reorganisation energy: lambda = sum c^2 / (2 w^2)
    assert np.isclose(np.sum(c ** 2 / (2 * w ** 2)), 0.625 * 100 / 101)


def test_wigner_widths_classical_and_quantum_limits():
    w, m = 2.0, 1.0
    hot = N.PositionHarmonicWigner(w, 1e-4, m)
    assert np.isclose(hot.sigma ** 2, 1 / (1e-4 * m * w ** 2), rtol=1e-6)      # kT / (m w^2)
    cold = N.PositionHarmonicWigner(w, 1e4, m)
    assert np.isclose(cold.sigma ** 2, 1 / (2 * m * w))                        # zero-point width
    assert np.isclose(N.VelocityHarmonicWigner(w, 1e4, m).sigma ** 2, w / (2 * m))


def test_distribution_moments_column_major():
    pos = np.array([[N.Normal(1.0, 0.1), N.Normal(2.0, 0.2), 3.0]])
    d = N.DynamicalDistribution(0.5, pos, (1, 3))
    mr, sr, mv, sv = d.moments()
    assert np.allclose(mr, [1, 2, 3]) and np.allclose(sr, [0.1, 0.2, 0]) and np.allclose(mv, 0.5) and np.allclose(sv, 0)
    p = d * N.PureState(1, N.Adiabatic())
    assert isinstance(p, N.distributions.ProductDistribution) and not p.electronic.diabatic
    assert N.PureState(2).diabatic  # PureState defaults to the diabatic basis


def test_simulation_syntax_and_rescaling_validation():
    sim = N.Simulation[N.FSSH](N.Atoms(2000), N.TullyModelOne(), rescaling="vinversion")
    assert sim.size() == (1, 1) and sim.method.rescaling_code == 1
    with pytest.raises(ValueError, match="rescaling"):
        N.Simulation[N.FSSH](N.Atoms(2000), N.TullyModelOne(), rescaling="sideways")
    sb = N.SpinBoson(N.DebyeSpectralDensity(5.0, 0.625), 100, 1.0, 1.0)
    sim = N.Simulation[N.Ehrenfest](N.Atoms([1.0] * 100), sb)
    assert sim.size() == (1, 100)
    with pytest.raises(ValueError):
        N.Simulation[N.Ehrenfest](N.Atoms([1.0] * 10), sb)


def test_run_dynamics_rejects_before_touching_the_device():
    sim = N.Simulation[N.FSSH](N.Atoms(2000), N.TullyModelOne())
    u = N.DynamicsVariables(sim, [[0.01]], [[-5.0]], N.PureState(1, N.Adiabatic()))
    with pytest.raises(ValueError, match="EnsembleMI355X"):
        N.run_dynamics(sim, (0.0, 10.0), u, output=N.OutputSurfaceHops, ensemble_algorithm=object())
    with pytest.raises(ValueError, match="saveat"):
        N.run_dynamics(sim, (0.0, 10.0), u, output=N.OutputSurfaceHops, dt=1.0, saveat=0.5)
    with pytest.raises(ValueError, match="TerminatingCallback"):
        N.run_dynamics(sim, (0.0, 10.0), u, output=N.OutputSurfaceHops, callback=lambda u, t, i: True)
    rp = N.RingPolymerSimulation[N.Classical](N.Atoms(2000), N.DoubleWellGaussian(nstates=1), 8)
    with pytest.raises(ValueError, match="does not support"):
        N.run_dynamics(rp, (0.0, 10.0), u, output=N.OutputSurfaceHops)
