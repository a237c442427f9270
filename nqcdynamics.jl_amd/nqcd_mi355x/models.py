"""Model Hamiltonians understood by the device kernels (NQCModels names and defaults).

Each model carries the kind code and parameter vector that nqcd_mi355x_set_model
takes.  Parameters follow the published definitions: Tully, J. Chem. Phys. 93,
1061 (1990) for the three Tully models; the spin-boson discretisations of
docs/src/NQCModels/systembathmodels.md:21-27,54-66,89-96.
"""
import numpy as np

from . import _lib as L


class Model:
    nstates = 2
    ndofs = 1
    kind = None

    def params(self):
        raise NotImplementedError

    def __repr__(self):
        return f"{type(self).__name__}({', '.join(f'{k}={v!r}' for k, v in vars(self).items())})"


class TullyModelOne(Model):
    """Single avoided crossing: V11 = A(1 - e^{-B|x|}) sign x, V22 = -V11, V12 = C e^{-D x^2}."""
    kind = L.MODEL_TULLY1

    def __init__(self, A=0.01, B=1.6, C=0.005, D=1.0):
        self.A, self.B, self.C, self.D = A, B, C, D

    def params(self):
        return [self.A, self.B, self.C, self.D]


class TullyModelTwo(Model):
    """Dual avoided crossing: V11 = 0, V22 = -A e^{-B x^2} + E0, V12 = C e^{-D x^2}."""
    kind = L.MODEL_TULLY2

    def __init__(self, A=0.1, B=0.28, C=0.015, D=0.06, E0=0.05):
        self.A, self.B, self.C, self.D, self.E0 = A, B, C, D, E0

    def params(self):
        return [self.A, self.B, self.C, self.D, self.E0]


class TullyModelThree(Model):
    """Extended coupling with reflection: V11 = A, V22 = -A, V12 = B e^{Cx} (x<0), B(2 - e^{-Cx}) (x>=0)."""
    kind = L.MODEL_TULLY3

    def __init__(self, A=6e-4, B=0.1, C=0.9):
        self.A, self.B, self.C = A, B, C

    def params(self):
        return [self.A, self.B, self.C]


class DoubleWellGaussian(Model):
    """Quartic double well V11 = a x^4 - b x^2, V22 = V11 + offset, Gaussian coupling V12 = C e^{-D x^2}.
    With nstates=1 only V11 is used (ring-polymer classical dynamics)."""
    kind = L.MODEL_DW_GAUSSIAN

    def __init__(self, a=1 / 1200, b=1 / 40, offset=0.01, C=0.005, D=1.0, nstates=2):
        self.a, self.b, self.offset, self.C, self.D = a, b, offset, C, D
        self.nstates = nstates

    def params(self):
        return [self.a, self.b, self.offset, self.C, self.D]


class OhmicSpectralDensity:
    """J(w) = (pi/2) alpha w e^{-w/wc}; w_j = -wc ln(1 - j/(N+1)), c_j = sqrt(alpha wc/(N+1)) w_j."""

    def __init__(self, wc, alpha):
        self.wc, self.alpha = wc, alpha

    def discretise(self, N):
        j = np.arange(1, N + 1)
        w = -self.wc * np.log(1 - j / (N + 1))
        return w, np.sqrt(self.alpha * self.wc / (N + 1)) * w


class DebyeSpectralDensity:
    """J(w) = 2 lambda wc w / (wc^2 + w^2); w_j = wc tan(pi/2 (1 - j/(N+1))), c_j = sqrt(2 lambda/(N+1)) w_j."""

    def __init__(self, wc, lam):
        self.wc, self.lam = wc, lam

    def discretise(self, N):
        j = np.arange(1, N + 1)
        w = self.wc * np.tan(np.pi / 2 * (1 - j / (N + 1)))
        return w, np.sqrt(2 * self.lam / (N + 1)) * w


class SpinBoson(Model):
    """[[eps + c.R, Delta], [Delta, -eps - c.R]] + R^T W^2 R / 2 with unit bath masses;
    bath positions laid out (1, N) as in the reference."""
    kind = L.MODEL_SPIN_BOSON

    def __init__(self, density, N, eps, Delta):
        self.density, self.N, self.eps, self.Delta = density, N, eps, Delta
        self.omega, self.c = density.discretise(N)
        self.ndofs = 1

    @property
    def natoms(self):
        return self.N

    def params(self):
        return np.concatenate([[self.eps, self.Delta], self.omega, self.c])


def nstates(model):
    return model.nstates


class Harmonic(Model):
    """Single-state harmonic well V = m w^2 (x - r0)^2 / 2 (ring-polymer tests)."""
    kind = L.MODEL_HARMONIC
    nstates = 1

    def __init__(self, m=1.0, omega=1.0, r0=0.0):
        self.m, self.omega, self.r0 = m, omega, r0

    def params(self):
        return [self.m, self.omega, self.r0]


class MiaoSubotnik:
    """Impurity diabats U0 = m w^2 x^2 / 2 and U1 = m w^2 (x - g)^2 / 2 + dG with hybridisation Gamma
    (NQCModels' MiaoSubotnik defaults)."""

    def __init__(self, m=2000.0, omega=2e-4, g=20.6097, Gamma=6.4e-3, dG=-3.8e-3):
        self.m, self.omega, self.g, self.Gamma, self.dG = m, omega, g, Gamma, dG


class TrapezoidalRule:
    """Wide-band bath of M evenly spaced levels eps_n = a + (n-1)(b-a)/(M-1) with couplings
    sqrt(Gamma/2pi) sqrt((b-a)/(M-1)) (docs/src/NQCModels/systembathmodels.md, TrapezoidalRule)."""

    def __init__(self, M, bandmin, bandmax):
        if M < 2:
            raise ValueError("TrapezoidalRule needs M >= 2")
        self.M, self.bandmin, self.bandmax = int(M), float(bandmin), float(bandmax)

    def energies(self):
        return self.bandmin + np.arange(self.M) * (self.bandmax - self.bandmin) / (self.M - 1)


class AndersonHolstein(Model):
    """Newns-Anderson impurity model: the impurity of `model` coupled to the discretised `bath`.
    nstates = M + 1 and, unless given, M // 2 electrons (half filling: the count the reference's
    IESH / EhrenfestNA tests use with every Fermi level they set, test/Dynamics/iesh.jl:18-30,
    ehrenfest_na.jl:12-19).  A Fermi level shifts every orbital energy alike, which changes neither
    the forces nor the populations of the dynamics."""
    kind = L.MODEL_NEWNS_ANDERSON

    def __init__(self, model, bath, fermi_level=0.0, nelectrons=None):
        if not isinstance(model, MiaoSubotnik):
            raise ValueError("EnsembleMI355X provides the MiaoSubotnik impurity")
        self.model, self.bath, self.fermi_level = model, bath, float(fermi_level)
        self.nstates = bath.M + 1
        self.nelectrons = bath.M // 2 if nelectrons is None else int(nelectrons)

    def params(self):
        m = self.model
        return [m.m, m.omega, m.g, m.dG, m.Gamma, self.bath.bandmin, self.bath.bandmax, self.bath.M]


def nelectrons(model):
    return model.nelectrons


class MorseN(Model):
    """n-state Morse model of one coordinate (Coker and Xiao, J. Chem. Phys. 102, 496 (1995); NQCModels'
    ThreeStateMorse is n = 3): V_ii = D_i (1 - e^{-beta_i (x - r_i)})^2 + c_i, V_ij = A_ij e^{-a_ij (x - r_ij)^2}.
    `states` is a list of (D, beta, r, c) per state and `couplings` a dict {(i, j): (A, a, r)} over i < j
    (0-based; missing pairs are uncoupled)."""
    kind = L.MODEL_MORSE_N

    def __init__(self, states, couplings=None):
        self.states = [tuple(float(x) for x in st) for st in states]
        self.nstates = len(self.states)
        if not 2 <= self.nstates <= 8:
            raise ValueError("MorseN takes 2 to 8 states")
        self.couplings = {tuple(k): tuple(float(x) for x in v) for k, v in (couplings or {}).items()}

    @classmethod
    def from_table(cls, n, table):
        t = np.asarray(table, dtype=np.float64)
        st = t[:4 * n].reshape(n, 4)
        pairs = t[4 * n:].reshape(-1, 3)
        cp, k = {}, 0
        for i in range(n):
            for j in range(i + 1, n):
                cp[(i, j)] = tuple(pairs[k]); k += 1
        return cls(st, cp)

    def params(self):
        n = self.nstates
        out = [n] + [x for st in self.states for x in st]
        for i in range(n):
            for j in range(i + 1, n):
                out += list(self.couplings.get((i, j), (0.0, 1.0, 0.0)))
        return out


class AndersonHolsteinOneMode(Model):
    """One-mode Anderson-Holstein impurity of the CME / BCME tests (test/Dynamics/cme.jl:27-55): V11 = hw x^2/2,
    V22 = V11 + Ed + sqrt(2) g x, V12 = sqrt(Gamma / 2 pi); Fermi level of the metal `fermi_level`."""
    kind = L.MODEL_ANDERSON_HOLSTEIN_1

    def __init__(self, hw, Ed, g, Gamma, fermi_level=0.0):
        self.hw, self.Ed, self.g, self.Gamma, self.fermi_level = float(hw), float(Ed), float(g), float(Gamma), float(fermi_level)

    def params(self):
        return [self.hw, self.Ed, np.sqrt(2) * self.g, np.sqrt(self.Gamma / (2 * np.pi))]


class DiatomSurface(Model):
    """Diatomic molecule above a surface, six coordinates (ndofs = 3, natoms = 2), two diabats:
    bond b = |R1 - R2|, height Z = mu z1 + (1 - mu) z2,
      V11 = Dn (1 - e^{-an (b - bn)})^2 + Aw e^{-bw Z}              (neutral molecule, repulsive wall)
      V22 = Di (1 - e^{-ai (b - bi)})^2 + Ds (1 - e^{-as (Z - Zs)})^2 + Delta   (ion bound to the surface)
      V12 = C e^{-g (Z - Zc)^2}.
    `seed` jitters every parameter but mu by up to 5% (a reproducible synthetic model family)."""
    kind = L.MODEL_DIATOM_SURFACE
    ndofs = 3
    natoms = 2
    BASE = (0.24, 1.1, 2.2, 2.0, 1.2, 0.20, 1.0, 2.35, 0.12, 0.8, 3.0, 0.0, 0.003, 1.0, 3.75, 0.5)
    NAMES = ("Dn", "an", "bn", "Aw", "bw", "Di", "ai", "bi", "Ds", "as", "Zs", "Delta", "C", "g", "Zc", "mu")

    def __init__(self, params=None, seed=None):
        p = np.array(self.BASE if params is None else params, dtype=np.float64)
        if seed is not None:
            jit = 1.0 + 0.05 * (2.0 * np.random.default_rng(seed).random(len(p)) - 1.0)
            jit[-1] = 1.0
            p = p * jit
        self.p = p

    def params(self):
        return [float(x) for x in self.p]

    def potential(self, X):
        """(n, 6) coordinates -> (n, 3) diabatic elements (V11, V22, V12)."""
        X = np.atleast_2d(np.asarray(X, dtype=np.float64))
        Dn, an, bn, Aw, bw, Di, ai, bi, Ds, a_s, Zs, Delta, C, g, Zc, mu = self.p
        b = np.linalg.norm(X[:, :3] - X[:, 3:], axis=1)
        Z = mu * X[:, 2] + (1 - mu) * X[:, 5]
        morse = lambda D, a, q0, x: D * (1 - np.exp(-a * (x - q0))) ** 2
        v11 = morse(Dn, an, bn, b) + Aw * np.exp(-bw * Z)
        v22 = morse(Di, ai, bi, b) + morse(Ds, a_s, Zs, Z) + Delta
        v12 = C * np.exp(-g * (Z - Zc) ** 2)
        return np.stack([v11, v22, v12], axis=1)


class MLPSurrogate(Model):
    """The MLP surrogate (6 -> 128 -> 128 -> 3, two ensemble members) standing in for the diabatic
    Hamiltonian of `analytic`; wherever the members disagree by more than `threshold` (Hartree) on
    an element of H the analytic model is used for that trajectory step."""
    kind = L.MODEL_MLP_SURROGATE
    ndofs = 3
    natoms = 2

    def __init__(self, analytic, members, norm, threshold=2e-3):
        self.analytic, self.members, self.norm, self.threshold = analytic, [np.asarray(m, np.float32) for m in members], \
            np.asarray(norm, np.float32), float(threshold)

    def params(self):
        return [self.threshold] + self.analytic.params()


def train_surrogate(analytic, X, Y=None, threshold=2e-3, device=0, **train_kw):
    """Train the two-member surrogate of `analytic` on configurations X (n, 6) (labels Y = the
    analytic diabatic elements unless given) with the device's BF16 training kernels."""
    Y = analytic.potential(X) if Y is None else Y
    E = L.Ensemble(device)
    try:
        hist = E.surrogate_train(X, Y, **train_kw)
        members, norm = [], None
        for k in range(2):
            p, norm = E.surrogate_get_params(k)
            members.append(p)
    finally:
        E.close()
    return MLPSurrogate(analytic, members, norm, threshold), hist
