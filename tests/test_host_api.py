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
    na = N.AndersonHolstein(N.MiaoSubotnik(), N.TrapezoidalRule(40, -0.0192, 0.0192))
    cl = N.Simulation[N.Classical](N.Atoms(2000), N.TullyModelOne())
    with pytest.raises(ValueError, match="does not support"):
        N.run_dynamics(cl, (0.0, 10.0), u, output=N.OutputSurfaceHops)
    rp = N.RingPolymerSimulation[N.AdiabaticIESH](N.Atoms(2000), na, 8, temperature=9.5e-4)
    with pytest.raises(ValueError, match="bead configurations"):
        N.run_dynamics(rp, (0.0, 10.0), N.DynamicalDistribution(N.Normal(0.0, 1e-2), N.Normal(8.0, 1.0), (1, 1, 8)),
                       output=N.OutputDiabaticPopulation)
    ie = N.Simulation[N.AdiabaticIESH](N.Atoms(2000), na)
    with pytest.raises(ValueError, match="termination"):
        N.run_dynamics(ie, (0.0, 10.0), u, output=N.OutputSurfaceHops, callback=N.TerminatingCallback(1, 10.0))


def test_newns_anderson_model_and_iesh_options():
    """AndersonHolstein(MiaoSubotnik(), TrapezoidalRule(M, -W, W)): M + 1 states, M/2 electrons for a
    symmetric band (test/Dynamics/iesh.jl:18-30); method keywords of iesh.jl:72-87."""
    W = 6 * 6.4e-3 / 2
    bath = N.TrapezoidalRule(30, -W, W)
    assert np.isclose(bath.energies()[0], -W) and np.isclose(bath.energies()[-1], W)
    na = N.AndersonHolstein(N.MiaoSubotnik(), bath)
    assert N.nstates(na) == 31 and N.nelectrons(na) == 15
    assert N.nelectrons(N.AndersonHolstein(N.MiaoSubotnik(), bath, fermi_level=0.001)) == 15
    assert na.params() == [2000.0, 2e-4, 20.6097, -3.8e-3, 6.4e-3, -W, W, 30]
    sim = N.Simulation[N.AdiabaticIESH](N.Atoms(2000), na, rescaling="vinversion", estimate_probability=False,
                                        decoherence=N.DecoherenceCorrectionEDC(0.2))
    assert sim.method.rescaling_code == 1 and not sim.method.estimate_probability and sim.method.decoherence.C == 0.2
    with pytest.raises(ValueError):
        N.Simulation[N.AdiabaticIESH](N.Atoms(2000), na, disable_hopping=True)
    with pytest.raises(ValueError):
        N.Simulation[N.AdiabaticIESH](N.Atoms(2000), na, rescaling="other")


def test_fermi_dirac_state_checks():
    """FermiDiracState initial conditions: beta = 1/kT (inf at 0 K); the model's and the distribution's
    Fermi levels must agree (test/Dynamics/iesh.jl:43-46); IESH / EhrenfestNA only; the Diabatic basis for IESH only;
    available states within the system (DynamicsUtils.jl:211-214).  All raise before the device opens."""
    assert N.FermiDiracState(0.0, 0.0).beta == np.inf
    assert np.isclose(N.FermiDiracState(0.0, 1e-3).beta, 1e3)
    with pytest.raises(ValueError):
        N.FermiDiracState(0.0, -1.0)
    W = 6 * 6.4e-3 / 2
    na = N.AndersonHolstein(N.MiaoSubotnik(), N.TrapezoidalRule(30, -W, W), fermi_level=0.1 / 27.211386)
    ie = N.Simulation[N.AdiabaticIESH](N.Atoms(2000), na)
    nuc = N.DynamicalDistribution(N.Normal(0.0, 1e-2), N.Normal(8.0, 1.0), (1, 1))
    with pytest.raises(ValueError, match="Fermi level"):
        N.run_dynamics(ie, (0.0, 10.0), nuc * N.FermiDiracState(0.0, 9.5e-4), output=N.OutputDiabaticPopulation)
    ena = N.Simulation[N.EhrenfestNA](N.Atoms(2000), na)
    with pytest.raises(ValueError, match="Diabatic"):
        N.run_dynamics(ena, (0.0, 10.0), nuc * N.FermiDiracState(na.fermi_level, 9.5e-4, basis=N.Diabatic()),
                       output=N.OutputDiabaticPopulation)
    with pytest.raises(ValueError, match="only 31 states"):
        N.run_dynamics(ie, (0.0, 10.0), nuc * N.FermiDiracState(na.fermi_level, 9.5e-4, available_states=range(1, 33)),
                       output=N.OutputDiabaticPopulation)
    sh = N.Simulation[N.FSSH](N.Atoms(2000), N.TullyModelOne())
    with pytest.raises(ValueError, match="IESH"):
        N.run_dynamics(sh, (0.0, 10.0), nuc * N.FermiDiracState(0.0, 9.5e-4), output=N.OutputDiabaticPopulation)


def test_terminating_callback_predicates():
    """The device-evaluated TerminatingCallback predicates map onto the C ABI's kinds."""
    from nqcd_mi355x import _lib as L
    assert N.TerminatingCallback(1, 10.0).predicate == "abs_gt"
    assert N.TerminatingCallback.PREDICATES == {"abs_gt": L.TERM_ABS_GT, "gt": L.TERM_GT, "lt": L.TERM_LT}
    assert (L.TERM_ABS_GT, L.TERM_GT, L.TERM_LT) == (1, 2, 3)
    with pytest.raises(ValueError, match="predicate"):
        N.TerminatingCallback(1, 10.0, predicate="ge")


def test_save_schedule_forms():
    """saveat as SciML takes it: None -> every step; a number -> t0:saveat:t1; an evenly spaced sequence from
    t0 -> its own times; anything else raises."""
    from nqcd_mi355x.ensembles import save_schedule
    k, n, t = save_schedule(None, 0.0, 10.0, 0.5)
    assert (k, n) == (1, 21) and np.allclose(t, np.arange(21) * 0.5)
    k, n, t = save_schedule(2.0, 0.0, 10.0, 0.5)
    assert (k, n) == (4, 6) and np.allclose(t, [0, 2, 4, 6, 8, 10])
    k, n, t = save_schedule(np.arange(0.0, 6.1, 1.5), 0.0, 10.0, 0.5)
    assert (k, n) == (3, 5) and np.allclose(t, [0, 1.5, 3, 4.5, 6])
    k, n, t = save_schedule([0.0], 0.0, 10.0, 0.5)
    assert (k, n) == (0, 1)
    for bad in (0.3, 3.0, [0.0, 1.0, 3.0], [1.0, 2.0], np.arange(0.0, 12.0, 1.0)):
        with pytest.raises(ValueError):
            save_schedule(bad, 0.0, 10.0, 0.5)
