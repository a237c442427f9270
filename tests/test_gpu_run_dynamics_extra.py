"""run_dynamics (the mirrored reference API) for IESH on the Newns-Anderson model and for ring-polymer
surface hopping, against the FP64 oracle on the same uploaded initial conditions."""
import numpy as np
import pytest

import nqcd_mi355x as N

pytestmark = pytest.mark.gpu


def test_iesh_run_dynamics_mean_populations(oracle, key):
    M, ntraj, dt, nsteps = 40, 48, 2.0, 200
    W = 6 * 6.4e-3 / 2
    model = N.AndersonHolstein(N.MiaoSubotnik(), N.TrapezoidalRule(M, -W, W))
    sim = N.Simulation[N.AdiabaticIESH](N.Atoms(2000), model)
    rng = np.random.default_rng(5)
    r0 = rng.normal(8.0, 1.0, size=ntraj)
    v0 = rng.normal(0.0, 1e-2, size=ntraj)
    dist = N.DynamicalDistribution([np.array([[x]]) for x in v0], [np.array([[x]]) for x in r0], (1, 1))
    res = N.run_dynamics(sim, (0.0, nsteps * dt), dist, trajectories=ntraj, dt=dt, saveat=10 * dt,
                         output=(N.OutputDiabaticPopulation, N.OutputAdiabaticPopulation),
                         reduction=N.MeanReduction(), ensemble_algorithm=N.EnsembleMI355X(seed=key))
    md = oracle.na_model(M=M, W=W)
    occ0 = np.tile(np.arange(M // 2), (ntraj, 1))
    c = oracle.iesh_run_ensemble(0, md, M // 2, 2000.0, r0, v0, occ0, dt=dt, nsteps=nsteps, key=key, save_every=10)
    di = np.array(res["OutputDiabaticPopulation"])
    assert di.shape == (nsteps // 10 + 1, M + 1)
    assert np.allclose(di, c["pop_di"].mean(axis=0), rtol=1e-9, atol=1e-12)
    ad = np.array(res["OutputAdiabaticPopulation"])
    final = np.zeros(M + 1)
    for o in c["occ"]:
        final[o] += 1.0 / ntraj
    assert np.allclose(ad[-1], final, atol=1e-12)
    assert np.allclose(res["Time"], np.arange(nsteps // 10 + 1) * 10 * dt)


@pytest.mark.parametrize("method,basis", [("AdiabaticIESH", "Adiabatic"), ("EhrenfestNA", "Adiabatic"),
                                          ("AdiabaticIESH", "Diabatic")])
def test_run_dynamics_fermi_dirac_initial_state(oracle, key, method, basis):
    """DynamicalDistribution * FermiDiracState(fermi_level, kT[, Diabatic()]): thermal occupations (and
    for the Diabatic form the orbitals) drawn on the device at each trajectory's position
    (iesh.jl:99-183, ehrenfest_na.jl:43-70), then the dynamics; mean diabatic populations against the
    oracle run from the oracle's occupations of the same streams."""
    M, ntraj, dt, nsteps = 40, 48, 2.0, 100
    W = 6 * 6.4e-3 / 2
    kT = 3.166811563e-6 * 300.0
    model = N.AndersonHolstein(N.MiaoSubotnik(), N.TrapezoidalRule(M, -W, W))
    sim = N.Simulation[getattr(N, method)](N.Atoms(2000), model)
    rng = np.random.default_rng(8)
    r0 = rng.normal(6.0, 1.0, size=ntraj)
    v0 = rng.normal(0.0, 1e-2, size=ntraj)
    dist = N.DynamicalDistribution([np.array([[x]]) for x in v0], [np.array([[x]]) for x in r0], (1, 1)) * \
        N.FermiDiracState(0.0, kT, basis=getattr(N, basis)())
    res = N.run_dynamics(sim, (0.0, nsteps * dt), dist, trajectories=ntraj, dt=dt, saveat=10 * dt,
                         output=N.OutputDiabaticPopulation, reduction=N.MeanReduction(),
                         ensemble_algorithm=N.EnsembleMI355X(seed=key))
    md = oracle.na_model(M=M, W=W)
    psi0 = None
    if basis == "Adiabatic":
        occ0 = np.stack([oracle.fermi_dirac_occupations(oracle.na_eigen(md, x)[0], M // 2, 1 / kT, t, key)
                         for t, x in enumerate(r0)])
        assert not np.all(occ0 == np.arange(M // 2))
    else:
        occ0, psi0 = [], []
        for t, x in enumerate(r0):
            U = oracle.na_eigen(md, x)[1]
            o, p = oracle.fermi_dirac_diabatic(np.diag(oracle.na_matrix(md, x)), U * np.sign(U[0]), M // 2, 1 / kT, t, key)
            occ0.append(o); psi0.append(p.astype(complex))
        occ0, psi0 = np.stack(occ0), np.stack(psi0)
    c = oracle.iesh_run_ensemble(0 if method == "AdiabaticIESH" else 1, md, M // 2, 2000.0, r0, v0, occ0, psi0=psi0,
                                 dt=dt, nsteps=nsteps, key=key, save_every=10)
    assert np.allclose(np.array(res["OutputDiabaticPopulation"]), c["pop_di"].mean(axis=0), rtol=1e-9, atol=1e-12)


def test_rpsh_run_dynamics_centroid_and_populations(oracle, key):
    P, ntraj, nsteps = 8, 64, 400
    DW2 = [1 / 1200, 1 / 40, 0.01, 0.005, 1.0]
    model = N.DoubleWellGaussian(*DW2)
    sim = N.RingPolymerSimulation[N.FSSH](N.Atoms(1836.0), model, P, temperature=1 / 8)
    rng = np.random.default_rng(6)
    r0 = rng.normal(-2.0, 0.05, size=(ntraj, 1)) + rng.normal(0, 0.05, size=(ntraj, P))
    v0 = rng.uniform(0.011, 0.018, size=(ntraj, 1)) + rng.normal(0, 1e-4, size=(ntraj, P))
    dist = N.DynamicalDistribution([v.reshape(1, 1, P) for v in v0], [r.reshape(1, 1, P) for r in r0], (1, 1, P))
    res = N.run_dynamics(sim, (0.0, float(nsteps)), dist * N.PureState(1, N.Adiabatic()), trajectories=ntraj, dt=1.0,
                         saveat=10.0, output=(N.OutputCentroidPosition, N.OutputDiabaticPopulation),
                         reduction=N.MeanReduction(), ensemble_algorithm=N.EnsembleMI355X(seed=key))
    m = oracle.make_model(oracle.MODEL_DW_GAUSSIAN, 2, 1, DW2)
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    c = oracle.rp_run_ensemble(oracle.RP_FSSH, m, 1836.0, r0, v0, sig0, np.zeros(ntraj, np.int32), dt=1.0,
                               nsteps=nsteps, temperature=1 / 8, save_every=10, nsave=nsteps // 10 + 1, key=key)
    assert c["nhops"].sum() > 0
    assert np.allclose(res["OutputCentroidPosition"], c["centroid"].mean(axis=0), rtol=1e-9, atol=1e-11)
    assert np.allclose(np.array(res["OutputDiabaticPopulation"]), c["pop_di"].mean(axis=0), rtol=1e-9, atol=1e-11)


def test_ring_polymer_iesh_run_dynamics(oracle, key):
    """RingPolymerSimulation{AdiabaticIESH}(atoms, model, n_beads; temperature) through run_dynamics
    (bead configurations from a list distribution, BCBWavefunction selected as select_algorithm does,
    IntegrationAlgorithms.jl:99) against the oracle: mean populations and final beads."""
    M, P, ntraj, dt, nsteps = 40, 4, 24, 2.0, 150
    W = 6 * 6.4e-3 / 2
    temp = 3.166811563e-6 * 300.0
    model = N.AndersonHolstein(N.MiaoSubotnik(), N.TrapezoidalRule(M, -W, W))
    sim = N.RingPolymerSimulation[N.AdiabaticIESH](N.Atoms(2000), model, P, temperature=temp)
    rng = np.random.default_rng(9)
    rb = rng.normal(8.0, 1.0, (ntraj, 1)) + rng.normal(0, 0.05, (ntraj, P))
    vb = np.abs(rng.normal(0.0, 1e-2, (ntraj, 1))) + rng.normal(0, 1e-3, (ntraj, P))
    dist = N.DynamicalDistribution([v.reshape(1, 1, P) for v in vb], [r.reshape(1, 1, P) for r in rb], (1, 1, P))
    res = N.run_dynamics(sim, (0.0, nsteps * dt), dist, trajectories=ntraj, dt=dt, saveat=10 * dt,
                         output=N.OutputDiabaticPopulation, reduction=N.MeanReduction(),
                         ensemble_algorithm=N.EnsembleMI355X(seed=key))
    md = oracle.na_model(M=M, W=W)
    c = oracle.rp_iesh_run_ensemble(0, md, M // 2, 2000.0, P, temp, rb, vb, np.tile(np.arange(M // 2), (ntraj, 1)),
                                    dt=dt, nsteps=nsteps, key=key, save_every=10)
    assert np.allclose(np.array(res["OutputDiabaticPopulation"]), c["pop_di"].mean(axis=0), rtol=1e-9, atol=1e-12)
