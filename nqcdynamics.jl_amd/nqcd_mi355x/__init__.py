"""EnsembleMI355X host package.

Mirrors the reference's user-facing API for ensemble trajectory integration --
`run_dynamics(sim, tspan, distribution; output, reduction, ensemble_algorithm,
trajectories, dt, saveat, callback)`, `Simulation{Method}`, the NQCModels models,
NQCDistributions initial conditions, DynamicsOutputs and reductions -- with every
trajectory integrated by libnqcd_mi355x.so (hand-written gfx950 kernels behind a
C ABI).  There is no CPU fallback: the library must be built.
"""
from . import _lib  # noqa: F401
from .distributions import (Adiabatic, Diabatic, DynamicalDistribution, DynamicsVariables, FermiDiracState, Normal,  # noqa: F401
                            PositionHarmonicWigner, PureState, VelocityBoltzmann, VelocityHarmonicWigner)
from .ensembles import (CellBoundaryCallback, EnsembleMI355X, TerminateCellCallback, TerminatingCallback,  # noqa: F401
                        run_dynamics)
from .models import (AndersonHolstein, DebyeSpectralDensity, DiatomSurface, DoubleWellGaussian, Harmonic,  # noqa: F401
                     MiaoSubotnik, MLPSurrogate, MorseN, AndersonHolsteinOneMode, OhmicSpectralDensity, SpinBoson,
                     TrapezoidalRule, TullyModelOne, TullyModelThree, TullyModelTwo, nelectrons, nstates,
                     train_surrogate)
from .outputs import (FileReduction, MeanReduction, OutputAdiabaticPopulation, OutputCentroidPosition,  # noqa: F401
                      OutputDiabaticPopulation, OutputDiscreteState, OutputFinal, OutputFinalTime,
                      OutputKineticEnergy, OutputPosition, OutputStateResolvedScattering1D, OutputSurfaceHops,
                      OutputTotalEnergy, OutputVelocity, PopulationCorrelationFunction, SortByOutputReduction,
                      SortByTrajectoryReduction, SumReduction)
from .simulations import (AdiabaticIESH, Atoms, BCME, CME, Classical, DecoherenceCorrectionEDC, Ehrenfest,  # noqa: F401
                          EhrenfestNA, FSSH, InfiniteCell, PeriodicCell, RingPolymerSimulation, Simulation, SpinMappingW, eCMM, NRPMD,
                          masses)
