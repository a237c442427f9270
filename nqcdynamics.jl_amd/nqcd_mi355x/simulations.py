"""Simulation containers and dynamics methods (src/simulations.jl, src/DynamicsMethods).

`Simulation[FSSH](atoms, model, rescaling="standard")` mirrors the reference's
`Simulation{FSSH}(atoms, model; rescaling=:standard)` (fssh.jl:38-45); the
method objects carry the options the device kernels read.
"""
import numpy as np

from . import _lib as L

_ELEMENT_MASS = {"H": 1837.4715941070515, "D": 3671.4829, "He": 7296.2993, "C": 12.011 * 1822.888486}


class Atoms:
    """Atom masses (NQCBase.Atoms): Atoms(2000), Atoms([1.0]*100) or Atoms("H")."""

    def __init__(self, spec):
        if isinstance(spec, str):
            spec = [spec]
        if np.isscalar(spec):
            spec = [spec]
        masses = [(_ELEMENT_MASS[s] if isinstance(s, str) else float(s)) for s in spec]
        self.masses = np.asarray(masses, dtype=np.float64)

    def __len__(self):
        return len(self.masses)


class Method:
    method_code = None
    algorithm = None


class FSSH(Method):
    """Fewest-switches surface hopping (fssh.jl:23-45); rescaling :standard, :vinversion or :off."""
    method_code = L.METHOD_FSSH
    algorithm = L.ALG_BAB_TSIT5

    def __init__(self, rescaling="standard"):
        codes = {"standard": L.RESCALE_STANDARD, "vinversion": L.RESCALE_VINVERSION, "off": L.RESCALE_OFF}
        if rescaling not in codes:
            raise ValueError("This mode of rescaling is not implemented")  # surface_hopping.jl:244
        self.rescaling = rescaling
        self.rescaling_code = codes[rescaling]
        self.state = 1


class Ehrenfest(Method):
    """Mean-field Ehrenfest dynamics (ehrenfest.jl:27-41)."""
    method_code = L.METHOD_EHRENFEST
    algorithm = L.ALG_BAB_TSIT5
    rescaling_code = L.RESCALE_STANDARD


class Classical(Method):
    """Classical molecular dynamics (classical.jl:22): VelocityVerlet for a Simulation (classical.jl:86),
    BCB for ring polymers."""
    method_code = L.METHOD_CLASSICAL
    algorithm = L.ALG_VELOCITY_VERLET
    rescaling_code = L.RESCALE_STANDARD


class InfiniteCell:
    """No simulation cell (NQCBase's InfiniteCell): every atom is always inside, the cell callbacks never fire."""


class PeriodicCell:
    """NQCBase's PeriodicCell(vectors; periodicity): the lattice vectors are the columns of `vectors`
    (ndofs x ndofs); `periodicity` flags the periodic dimensions (all by default)."""

    def __init__(self, vectors, periodicity=None):
        a = np.asarray(vectors, dtype=np.float64)
        self.vectors = a.reshape(1, 1) if a.ndim == 0 else a
        n = self.vectors.shape[0]
        if self.vectors.shape != (n, n) or not 1 <= n <= 3:
            raise ValueError("PeriodicCell vectors must be a square matrix of size 1 to 3")
        self.periodicity = np.ones(n, bool) if periodicity is None else np.asarray(periodicity, bool)
        if self.periodicity.shape != (n,):
            raise ValueError("periodicity needs one flag per dimension")


class DecoherenceCorrectionEDC:
    """Granucci-Persico energy decoherence correction (decoherence_corrections.jl:10-14)."""

    def __init__(self, C=0.1):
        self.C = float(C)


class AdiabaticIESH(Method):
    """Independent electron surface hopping (iesh.jl:26-87): rescaling, estimate_probability,
    disable_hopping and decoherence keywords."""
    method_code = L.METHOD_IESH
    algorithm = L.ALG_VERLET_ELECTRONICS

    def __init__(self, rescaling="standard", estimate_probability=True, disable_hopping=False, decoherence=None):
        codes = {"standard": L.RESCALE_STANDARD, "vinversion": L.RESCALE_VINVERSION, "off": L.RESCALE_OFF}
        if rescaling not in codes:
            raise ValueError("This mode of rescaling is not implemented")
        if disable_hopping:
            raise ValueError("EnsembleMI355X does not provide disable_hopping=true")
        if decoherence is not None and not isinstance(decoherence, DecoherenceCorrectionEDC):
            raise ValueError("decoherence must be DecoherenceCorrectionEDC() or None")
        self.rescaling, self.rescaling_code = rescaling, codes[rescaling]
        self.estimate_probability = bool(estimate_probability)
        self.decoherence = decoherence


class EhrenfestNA(Method):
    """Ehrenfest dynamics of the many-electron Newns-Anderson wavefunction (ehrenfest_na.jl:6-41)."""
    method_code = L.METHOD_EHRENFEST_NA
    algorithm = L.ALG_VERLET_ELECTRONICS
    rescaling_code = L.RESCALE_STANDARD


class CME(Method):
    """Classical master equation surface hopping for Anderson-Holstein models (cme.jl:63-113); two states,
    integrated by Tsit5 at the fixed step dt (IntegrationAlgorithms.jl:68)."""
    method_code = L.METHOD_CME
    algorithm = L.ALG_TSIT5
    rescaling_code = L.RESCALE_OFF
    bandwidth = 0.0


class BCME(Method):
    """Broadened classical master equation (cme.jl:115-182): CME with the Lorentzian-broadened Fermi force
    over `bandwidth` around the Fermi level."""
    method_code = L.METHOD_BCME
    algorithm = L.ALG_TSIT5
    rescaling_code = L.RESCALE_OFF

    def __init__(self, bandwidth):
        self.bandwidth = float(bandwidth)
        if not self.bandwidth > 0:
            raise ValueError("bandwidth must be positive")


class SpinMappingW(Method):
    """Spin-mapping method (spin_mapping.jl:1-13) for n electronic states: gamma = 2 (sqrt(n + 1) - 1) / n,
    R^2 = 2 sqrt(n + 1); integrated by MInt (IntegrationAlgorithms.jl:79).  Simulation[SpinMappingW](atoms,
    model) takes n from the model, as Simulation{SpinMappingW} does (spin_mapping.jl:11-13)."""
    method_code = L.METHOD_SPIN_MAPPING
    algorithm = L.ALG_MINT
    rescaling_code = L.RESCALE_STANDARD

    def __init__(self, n_states=None):
        self.n_states = None
        if n_states is not None:
            self.bind(n_states)

    def bind(self, n_states):
        n = int(n_states)
        if n < 1:
            raise ValueError("n_states must be positive")
        self.n_states = n
        self.gamma = 2.0 / n * (np.sqrt(n + 1.0) - 1.0)
        self.R2 = 2.0 * np.sqrt(n + 1.0)


class eCMM(Method):
    """Extended classical mapping model (cmm.jl:1-24): Simulation[eCMM](atoms, model, gamma=0.0) mirrors
    Simulation{eCMM}(atoms, model; γ=0); integrated by MInt (IntegrationAlgorithms.jl:78)."""
    method_code = L.METHOD_ECMM
    algorithm = L.ALG_MINT
    rescaling_code = L.RESCALE_STANDARD

    def __init__(self, gamma=0.0):
        self.gamma = float(gamma)


class NRPMD(Method):
    """Nonadiabatic ring-polymer molecular dynamics (nrpmd.jl:1-47): RingPolymerSimulation[NRPMD](atoms, model,
    n_beads, gamma=0.5, temperature=T) mirrors RingPolymerSimulation{NRPMD}(atoms, model, n_beads; γ=0.5);
    integrated by RingPolymerMInt (IntegrationAlgorithms.jl:103)."""
    method_code = L.METHOD_NRPMD
    algorithm = L.ALG_RP_MINT
    rescaling_code = L.RESCALE_STANDARD

    def __init__(self, gamma=0.5):
        self.gamma = float(gamma)


class AbstractSimulation:
    def __init__(self, method, atoms, model, temperature=0.0, precision="fp64", cell=None, **method_kwargs):
        self.method = method(**method_kwargs) if isinstance(method, type) else method
        if isinstance(self.method, SpinMappingW) and self.method.n_states is None:
            self.method.bind(model.nstates)
        self.cell = InfiniteCell() if cell is None else cell
        if not isinstance(self.cell, (InfiniteCell, PeriodicCell)):
            raise ValueError("cell must be a PeriodicCell or an InfiniteCell")
        self.atoms = atoms
        self.model = model
        self.temperature = float(temperature)
        if precision not in ("fp64", "fp32_nuclei"):
            raise ValueError("precision must be 'fp64' or 'fp32_nuclei'")
        self.precision = precision
        natoms = getattr(model, "natoms", len(atoms))
        if natoms != len(atoms):
            raise ValueError(f"model has {natoms} atoms but Atoms has {len(atoms)}")

    @property
    def ndofs(self):
        return self.model.ndofs

    @property
    def natoms(self):
        return len(self.atoms)

    def size(self):
        return (self.ndofs, self.natoms)


class _SimulationMeta(type):
    def __getitem__(cls, method):
        def make(atoms, model, *args, **kw):
            return cls(method, atoms, model, *args, **kw)
        return make


class Simulation(AbstractSimulation, metaclass=_SimulationMeta):
    """Simulation{Method}(atoms, model; temperature) (simulations.jl:12-44)."""

    def __init__(self, method, atoms, model, **kw):
        super().__init__(method, atoms, model, **kw)
        self.nbeads = 1


class RingPolymerSimulation(AbstractSimulation, metaclass=_SimulationMeta):
    """RingPolymerSimulation{Method}(atoms, model, n_beads; temperature) (simulations.jl:46-73)."""

    def __init__(self, method, atoms, model, n_beads, **kw):
        super().__init__(method, atoms, model, **kw)
        self.nbeads = int(n_beads)

    def size(self):
        return (self.ndofs, self.natoms, self.nbeads)


def masses(sim):
    return sim.atoms.masses
