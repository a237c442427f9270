"""Initial-condition distributions (the NQCDistributions types the reference's
selections.jl:51-99 samples).

A `DynamicalDistribution(velocity, position, size)` combined with an electronic
`PureState(state, Adiabatic()|Diabatic())` gives a product distribution.  Gaussian
nuclear distributions (Normal, harmonic Wigner) are sampled on the device from
Philox stream 0; fixed arrays and lists of configurations are uploaded.
"""
import numpy as np


class Adiabatic:
    pass


class Diabatic:
    pass


class PureState:
    """Pure electronic state (1-based like the reference); default basis Diabatic."""

    def __init__(self, state, basis=None):
        self.state = int(state)
        self.basis = basis if basis is not None else Diabatic()

    @property
    def diabatic(self):
        return isinstance(self.basis, Diabatic) or self.basis is Diabatic


class FermiDiracState:
    """NQCDistributions' FermiDiracState(fermi_level, temperature; available_states) for IESH /
    EhrenfestNA: each trajectory's occupied adiabatic orbitals are drawn at its initial position by
    the Boltzmann-factor chain of sample_fermi_dirac_distribution (DynamicsUtils.jl:191-205), on the
    device (iesh.jl:99-128, ehrenfest_na.jl:43-70).  temperature is kT in Hartree (0 leaves the
    lowest orbitals occupied); available_states are 1-based like the reference's (None = all).  In
    the Diabatic basis (IESH, iesh.jl:138-183) the chain samples diabatic states and each electron's
    orbital is the corresponding diabatic state in the adiabatic basis."""

    def __init__(self, fermi_level, temperature, available_states=None, basis=None):
        self.fermi_level, self.temperature = float(fermi_level), float(temperature)
        if self.temperature < 0:
            raise ValueError("temperature must be >= 0")
        self.available_states = None if available_states is None else [int(s) for s in available_states]
        self.basis = basis if basis is not None else Adiabatic()

    @property
    def diabatic(self):
        return isinstance(self.basis, Diabatic) or self.basis is Diabatic

    @property
    def beta(self):
        return np.inf if self.temperature == 0 else 1.0 / self.temperature


class Normal:
    def __init__(self, mu, sigma):
        self.mu, self.sigma = float(mu), float(sigma)


def PositionHarmonicWigner(omega, beta, mass):
    """Wigner position distribution of a thermal harmonic oscillator: sigma^2 = 1/(2 m w tanh(beta w/2))."""
    return Normal(0.0, np.sqrt(1.0 / (2 * mass * omega * np.tanh(beta * omega / 2))))


def VelocityHarmonicWigner(omega, beta, mass):
    """Wigner velocity distribution: sigma^2 = w/(2 m tanh(beta w/2))."""
    return Normal(0.0, np.sqrt(omega / (2 * mass * np.tanh(beta * omega / 2))))


def VelocityBoltzmann(temperature, mass):
    """Maxwell-Boltzmann velocity of one degree of freedom (NQCDistributions' VelocityBoltzmann):
    Normal(0, sqrt(kT / m))."""
    return Normal(0.0, np.sqrt(float(temperature) / float(mass)))


def _moments(spec, shape):
    """(mean, sd) arrays of `shape` for a fixed value, Normal, or array of Normals/values."""
    if isinstance(spec, Normal):
        return np.full(shape, spec.mu), np.full(shape, spec.sigma)
    arr = np.asarray(spec, dtype=object) if not isinstance(spec, np.ndarray) else spec
    if arr.dtype == object:
        arr = np.broadcast_to(arr, shape)
        mu = np.zeros(shape); sd = np.zeros(shape)
        for idx, x in np.ndenumerate(arr):
            if isinstance(x, Normal):
                mu[idx], sd[idx] = x.mu, x.sigma
            else:
                mu[idx] = float(x)
        return mu, sd
    return np.broadcast_to(np.asarray(arr, dtype=np.float64), shape).copy(), np.zeros(shape)


class DynamicalDistribution:
    """DynamicalDistribution(velocity, position, size): velocity/position are a value, an
    array, a Normal, an array of Normals, or a list of per-trajectory arrays."""

    def __init__(self, velocity, position, size):
        self.velocity, self.position, self.size = velocity, position, tuple(size)

    def __mul__(self, electronic):
        return ProductDistribution(self, electronic)

    def is_list(self):
        return isinstance(self.position, (list, tuple)) and len(self.position) > 0 and \
            np.asarray(self.position[0]).shape == self.size

    def moments(self):
        """Flattened (mu_r, sd_r, mu_v, sd_v) in column-major (dofs, atoms) order."""
        mr, sr = _moments(self.position, self.size)
        mv, sv = _moments(self.velocity, self.size)
        f = lambda a: np.asarray(a).reshape(self.size).flatten(order="F")
        return f(mr), f(sr), f(mv), f(sv)


class ProductDistribution:
    def __init__(self, nuclear, electronic):
        self.nuclear, self.electronic = nuclear, electronic


def DynamicsVariables(sim, v, r, electronic=None):
    """A single initial condition (fssh.jl:47-65, ehrenfest.jl:43-48) in the reference's
    (ndofs, natoms) layout; the density is built on the device from `electronic`."""
    return {"v": np.asarray(v, dtype=np.float64).reshape(sim.size()),
            "r": np.asarray(r, dtype=np.float64).reshape(sim.size()),
            "electronic": electronic if electronic is not None else PureState(1, Adiabatic())}
