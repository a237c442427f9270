"""ctypes binding of the FP64 CPU oracle (oracle/nqcd_oracle.cpp).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libnqcd_oracle.so")

MODEL_TULLY1, MODEL_TULLY2, MODEL_TULLY3, MODEL_SPIN_BOSON, MODEL_DW_GAUSSIAN, MODEL_HARMONIC = range(6)
MODEL_DIATOM_SURFACE = 7
MODEL_MORSE_N = 9
MODEL_ANDERSON_HOLSTEIN_1 = 10
METHOD_FSSH, METHOD_EHRENFEST = 0, 1
RESCALE_STANDARD, RESCALE_VINVERSION, RESCALE_OFF = 0, 1, 2


class OModelDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nstates", C.c_int32), ("ndof", C.c_int32), ("nbath", C.c_int32),
                ("p", C.c_double * 16), ("omega", C.POINTER(C.c_double)), ("cbath", C.POINTER(C.c_double)),
                ("ext", C.POINTER(C.c_double))]


class OCell(C.Structure):
    """Simulation cell of the cell callbacks: mode 0 none, 1 CellBoundaryCallback, 2 TerminateCellCallback."""
    _fields_ = [("mode", C.c_int32), ("nd", C.c_int32), ("per", C.c_int32 * 3), ("pad", C.c_int32),
                ("a", C.c_double * 9), ("inv", C.c_double * 9)]


def make_cell(mode, vectors, periodic=None):
    """OCell of PeriodicCell(vectors; periodicity): lattice vectors as the columns of `vectors`."""
    c = OCell()
    if mode == 0:
        return c
    a = np.atleast_2d(np.asarray(vectors, dtype=np.float64))
    nd = a.shape[0]
    inv = np.linalg.inv(a)
    c.mode, c.nd = mode, nd
    for j in range(nd):
        c.per[j] = 1 if periodic is None else int(bool(periodic[j]))
        for k in range(nd):
            c.a[j + 3 * k], c.inv[j + 3 * k] = a[j, k], inv[j, k]
    return c


def cell_apply(cell, x):
    """The cell callback on one configuration (atoms of nd coordinates, coordinate fastest): returns
    (outside, x after the callback)."""
    x = np.array(x, dtype=np.float64).ravel()
    f = lib().oracle_cell_apply
    f.restype = C.c_int32
    out = f(C.byref(cell), C.c_int32(len(x)), _p(x))
    return bool(out), x


class ORunDesc(C.Structure):
    _fields_ = [("method", C.c_int32), ("nstates", C.c_int32), ("ndof", C.c_int32), ("rescaling", C.c_int32),
                ("dt", C.c_double), ("t0", C.c_double), ("nsteps", C.c_int64), ("term_x", C.c_double),
                ("save_every", C.c_int32), ("nsave", C.c_int32), ("key", C.c_uint64), ("traj_offset", C.c_int64),
                ("nsub", C.c_int32), ("term_dof", C.c_int32), ("term_kind", C.c_int32), ("pad", C.c_int32),
                ("cell", OCell)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def make_model(kind, nstates, ndof, params, omega=None, cbath=None):
    """Model descriptor; MODEL_MORSE_N takes its whole parameter table as params (any length)."""
    m = OModelDesc()
    m.kind, m.nstates, m.ndof = kind, nstates, ndof
    keep = []
    if kind == MODEL_MORSE_N:
        ext = np.ascontiguousarray(params, dtype=np.float64)
        if ext.size != 4 * nstates + 3 * nstates * (nstates - 1) // 2:
            raise ValueError("MORSE_N table: [D, beta, r0, c] per state, [A, a, r] per pair")
        keep.append(ext)
        m.ext = _p(ext)
        params = []
    for i, x in enumerate(params):
        m.p[i] = float(x)
    if omega is not None:
        om = np.ascontiguousarray(omega, dtype=np.float64)
        cb = np.ascontiguousarray(cbath, dtype=np.float64)
        keep += [om, cb]
        m.nbath = len(om)
        m.omega, m.cbath = _p(om), _p(cb)
    m._keep = keep  # keep the arrays alive with the descriptor
    return m


def philox(ctr, key):
    out = np.zeros(4, dtype=np.uint32)
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    lib().oracle_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(out, C.c_uint32))
    return out


def uniform(traj, step, stream, key):
    f = lib().oracle_uniform
    f.restype = C.c_double
    f.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64]
    return f(traj, step, stream, key)


def normal2(traj, step, stream, key):
    z = np.zeros(2)
    f = lib().oracle_normal2
    f.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(C.c_double)]
    f(traj, step, stream, key, _p(z))
    return z


def model_eval(m, r):
    n, D = m.nstates, m.ndof
    r = np.ascontiguousarray(r, dtype=np.float64)
    V = np.zeros((n, n)); dV = np.zeros((D, n, n)); V0 = C.c_double(0); dV0 = np.zeros(D)
    lib().oracle_model_eval(C.byref(m), _p(r), _p(V), _p(dV), C.byref(V0), _p(dV0))
    return V, dV, V0.value, dV0


def eigen_sym(H, Uprev=None):
    H = np.ascontiguousarray(H, dtype=np.float64)
    n = H.shape[0]
    E = np.zeros(n); U = np.zeros((n, n))
    up = None if Uprev is None else np.ascontiguousarray(Uprev, dtype=np.float64)
    lib().oracle_eigen_sym(C.c_int32(n), _p(H), _p(E), _p(U), _p(up))
    return E, U


def adiabatic_quantities(E, U, dV):
    n = len(E); D = dV.shape[0]
    G = np.zeros((D, n, n)); d = np.zeros((D, n, n))
    lib().oracle_adiabatic_quantities(C.c_int32(n), C.c_int32(D), _p(np.ascontiguousarray(E)), _p(np.ascontiguousarray(U)),
                                      _p(np.ascontiguousarray(dV)), _p(G), _p(d))
    return G, d


def propagate_density(sig, E0, W0, t0, E1, W1, t1, t, dt, nsub=5):
    n = sig.shape[0]
    sr = np.ascontiguousarray(sig.real, dtype=np.float64).copy()
    si = np.ascontiguousarray(sig.imag, dtype=np.float64).copy()
    f = lib().oracle_propagate_density
    f.argtypes = [C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                  C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double, C.c_int32]
    arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (E0, W0, E1, W1)]
    f(n, _p(sr), _p(si), _p(arr[0]), _p(arr[1]), t0, _p(arr[2]), _p(arr[3]), t1, t, dt, nsub)
    return sr + 1j * si


def fssh_probability(v, sig, s, d, dt):
    n = sig.shape[0]; D = len(v)
    prob = np.zeros(n)
    f = lib().oracle_fssh_probability
    f.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32,
                  C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double)]
    f(n, D, _p(np.ascontiguousarray(v, dtype=np.float64)), _p(np.ascontiguousarray(sig.real)), _p(np.ascontiguousarray(sig.imag)),
      s, _p(np.ascontiguousarray(d, dtype=np.float64)), dt, _p(prob))
    return prob


def fssh_select(prob, s, rnd):
    f = lib().oracle_fssh_select
    f.argtypes = [C.c_int32, C.POINTER(C.c_double), C.c_int32, C.c_double]
    f.restype = C.c_int32
    p = np.ascontiguousarray(prob, dtype=np.float64)
    return f(len(p), _p(p), s, rnd)


def rescale_velocity(v, mass, E, d, new_state, old_state, rescaling=RESCALE_STANDARD):
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    n = len(E); D = len(v)
    f = lib().oracle_rescale_velocity
    f.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                  C.POINTER(C.c_double), C.c_int32, C.c_int32]
    f.restype = C.c_int32
    ok = f(n, D, rescaling, _p(v), _p(np.ascontiguousarray(mass, dtype=np.float64)), _p(np.ascontiguousarray(E, dtype=np.float64)),
           _p(np.ascontiguousarray(d, dtype=np.float64)), new_state, old_state)
    return bool(ok), v


def sample_gaussian(ntraj, D, mu_r, sd_r, mu_v, sd_v, key, traj_offset=0, traj_ids=None):
    shape = (ntraj, D)
    arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), shape)) for a in (mu_r, sd_r, mu_v, sd_v)]
    r = np.zeros(shape); v = np.zeros(shape)
    f = lib().oracle_sample_gaussian
    f.argtypes = [C.c_int64, C.c_int64, C.c_int32] + [C.POINTER(C.c_double)] * 4 + [C.c_uint64, C.POINTER(C.c_double),
                                                                                    C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    tid = None if traj_ids is None else np.ascontiguousarray(traj_ids, dtype=np.int64)
    f(ntraj, traj_offset, D, *[_p(a) for a in arrs], key, _p(r), _p(v), _p(tid, C.c_int64))
    return r, v


def diabatic_to_adiabatic(m, r0, diabatic_state):
    ntraj = r0.shape[0]; n = m.nstates
    sr = np.zeros((ntraj, n, n)); si = np.zeros((ntraj, n, n))
    f = lib().oracle_diabatic_to_adiabatic
    f.argtypes = [C.POINTER(OModelDesc), C.c_int64, C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    f(C.byref(m), ntraj, _p(np.ascontiguousarray(r0, dtype=np.float64)), diabatic_state, _p(sr), _p(si))
    return sr + 1j * si


def run_ensemble(method, models, mass, r0, v0, sig0, state0=None, model_id=None, dt=1.0, nsteps=100, term_x=0.0,
                 save_every=0, nsave=1, key=0, traj_offset=0, rescaling=RESCALE_STANDARD, nthreads=None, t0=0.0, nsub=5,
                 traj_ids=None, term_dof=0, term_kind=0, energies=False, cell=None):
    """Run an ensemble on the CPU; returns a dict of per-trajectory results (with energies=True also
    the kinetic and potential energy of every save slot, ekin / epot (ntraj, nsave)); cell: an OCell."""
    if not isinstance(models, (list, tuple)):
        models = [models]
    ntraj, D = r0.shape
    n = models[0].nstates
    R = ORunDesc()
    R.method, R.nstates, R.ndof, R.rescaling = method, n, D, rescaling
    R.dt, R.t0, R.nsteps, R.term_x = dt, t0, nsteps, term_x
    R.save_every, R.nsave, R.key, R.traj_offset, R.nsub = save_every, nsave, key, traj_offset, nsub
    R.term_dof = term_dof
    R.term_kind = term_kind
    if cell is not None:
        R.cell = cell
    marr = (OModelDesc * len(models))(*models)
    mass = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, dtype=np.float64), (D,)))
    sr0 = np.ascontiguousarray(np.real(sig0), dtype=np.float64).reshape(ntraj, n, n)
    si0 = np.ascontiguousarray(np.imag(sig0), dtype=np.float64).reshape(ntraj, n, n)
    st0 = None if state0 is None else np.ascontiguousarray(state0, dtype=np.int32)
    mid = None if model_id is None else np.ascontiguousarray(model_id, dtype=np.int32)
    out = dict(r=np.zeros((ntraj, D)), v=np.zeros((ntraj, D)), sig_re=np.zeros((ntraj, n, n)), sig_im=np.zeros((ntraj, n, n)),
               state=np.zeros(ntraj, dtype=np.int32), steps=np.zeros(ntraj, dtype=np.int64), nhops=np.zeros(ntraj, dtype=np.int32),
               pop_ad=np.zeros((ntraj, nsave, n)), pop_di=np.zeros((ntraj, nsave, n)), save_state=np.zeros((ntraj, nsave), dtype=np.int32))
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    tid = None if traj_ids is None else np.ascontiguousarray(traj_ids, dtype=np.int64)
    if energies:
        if traj_ids is not None:
            raise ValueError("energies with traj_ids are not provided")
        out["ekin"], out["epot"] = np.zeros((ntraj, nsave)), np.zeros((ntraj, nsave))
        f = lib().oracle_run_ensemble_energies
        f.restype = C.c_int32
        f(C.byref(R), marr, _p(mass), C.c_int64(ntraj), _p(mid, C.c_int32),
          _p(np.ascontiguousarray(r0, dtype=np.float64)), _p(np.ascontiguousarray(v0, dtype=np.float64)), _p(sr0), _p(si0),
          _p(st0, C.c_int32), _p(out["r"]), _p(out["v"]), _p(out["sig_re"]), _p(out["sig_im"]), _p(out["state"], C.c_int32),
          _p(out["steps"], C.c_int64), _p(out["nhops"], C.c_int32), _p(out["pop_ad"]), _p(out["pop_di"]),
          _p(out["save_state"], C.c_int32), _p(out["ekin"]), _p(out["epot"]), C.c_int32(nthreads))
        return out
    f = lib().oracle_run_ensemble
    f.restype = C.c_int32
    f(C.byref(R), marr, _p(mass), C.c_int64(ntraj), _p(mid, C.c_int32),
      _p(np.ascontiguousarray(r0, dtype=np.float64)), _p(np.ascontiguousarray(v0, dtype=np.float64)), _p(sr0), _p(si0), _p(st0, C.c_int32),
      _p(out["r"]), _p(out["v"]), _p(out["sig_re"]), _p(out["sig_im"]), _p(out["state"], C.c_int32), _p(out["steps"], C.c_int64),
      _p(out["nhops"], C.c_int32), _p(out["pop_ad"]), _p(out["pop_di"]), _p(out["save_state"], C.c_int32), C.c_int32(nthreads),
      _p(tid, C.c_int64))
    return out


# --- ring polymers ---------------------------------------------------------------------------
RP_CLASSICAL, RP_FSSH, RP_EHRENFEST = 0, 1, 2


class ORPDesc(C.Structure):
    _fields_ = [("method", C.c_int32), ("nstates", C.c_int32), ("nbeads", C.c_int32), ("rescaling", C.c_int32),
                ("dt", C.c_double), ("t0", C.c_double), ("nsteps", C.c_int64), ("save_every", C.c_int32),
                ("nsave", C.c_int32), ("key", C.c_uint64), ("temperature", C.c_double), ("nsub", C.c_int32),
                ("pad", C.c_int32)]


def rp_normal_modes(P, wn, dt):
    U = np.zeros((P, P)); w = np.zeros(P); Cm = np.zeros((P, 2, 2))
    f = lib().oracle_rp_normal_modes
    f.argtypes = [C.c_int32, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    f(P, wn, dt, _p(U), _p(w), _p(Cm))
    return U, w, Cm


def rp_run_ensemble(method, model, mass, r0, v0, sig0=None, state0=None, dt=1.0, nsteps=100, temperature=1.0,
                    save_every=0, nsave=1, key=0, rescaling=RESCALE_STANDARD, nthreads=None, t0=0.0, nsub=5):
    """r0, v0: (ntraj, nbeads).  Returns final beads, density, states and per-save centroid/populations."""
    ntraj, P = r0.shape
    n = model.nstates
    R = ORPDesc()
    R.method, R.nstates, R.nbeads, R.rescaling = method, n, P, rescaling
    R.dt, R.t0, R.nsteps, R.save_every, R.nsave, R.key = dt, t0, nsteps, save_every, nsave, key
    R.temperature, R.nsub = temperature, nsub
    sr0 = si0 = None
    if sig0 is not None:
        sr0 = np.ascontiguousarray(np.real(sig0), dtype=np.float64)
        si0 = np.ascontiguousarray(np.imag(sig0), dtype=np.float64)
    st0 = None if state0 is None else np.ascontiguousarray(state0, dtype=np.int32)
    out = dict(r=np.zeros((ntraj, P)), v=np.zeros((ntraj, P)), sig_re=np.zeros((ntraj, n, n)), sig_im=np.zeros((ntraj, n, n)),
               state=np.zeros(ntraj, np.int32), nhops=np.zeros(ntraj, np.int32), centroid=np.zeros((ntraj, nsave)),
               pop_ad=np.zeros((ntraj, nsave, n)), pop_di=np.zeros((ntraj, nsave, n)), save_state=np.zeros((ntraj, nsave), np.int32))
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    f = lib().oracle_rp_run_ensemble
    f.restype = C.c_int32
    f(C.byref(R), C.byref(model), C.c_double(mass), C.c_int64(ntraj),
      _p(np.ascontiguousarray(r0, dtype=np.float64)), _p(np.ascontiguousarray(v0, dtype=np.float64)), _p(sr0), _p(si0),
      _p(st0, C.c_int32), _p(out["r"]), _p(out["v"]), _p(out["sig_re"]), _p(out["sig_im"]), _p(out["state"], C.c_int32),
      _p(out["nhops"], C.c_int32), _p(out["centroid"]), _p(out["pop_ad"]), _p(out["pop_di"]), _p(out["save_state"], C.c_int32),
      C.c_int32(nthreads))
    return out


class OPimcDesc(C.Structure):
    _fields_ = [("nbeads", C.c_int32), ("segment_length", C.c_int32), ("steps", C.c_int64), ("mass", C.c_double),
                ("temperature", C.c_double), ("delta", C.c_double), ("width", C.c_int32), ("reassess", C.c_int32),
                ("key", C.c_uint64)]


def pimc(model, x0, mass, temperature, delta, passes, segment=1.0, width=0, key=0, nthreads=None, reassess=1):
    """Path-integral Monte Carlo chains (MetropolisHastings.jl:131-148,170-258), one per row of x0 (ntraj, P):
    returns the final beads, bead velocities ~ N(0, sqrt(P T / m)) and accepted moves per chain."""
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    ntraj, P = x.shape
    D = OPimcDesc()
    D.nbeads, D.segment_length, D.steps = P, max(1, int(round(segment * P))), int(passes)
    D.mass, D.temperature, D.delta, D.width, D.key = mass, temperature, delta, width, key
    D.reassess = reassess
    v = np.zeros_like(x)
    acc = np.zeros(ntraj, np.int64)
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    f = lib().oracle_pimc
    f.restype = C.c_int32
    rc = f(C.byref(D), C.byref(model), C.c_int64(ntraj), _p(x), _p(v), _p(acc, C.c_int64), C.c_int32(nthreads))
    if rc != 0:
        raise ValueError("pimc: segment_length must be in 1..nbeads")
    return x, v, acc


class OCmeDesc(C.Structure):
    _fields_ = [("method", C.c_int32), ("save_every", C.c_int32), ("nsave", C.c_int32), ("pad", C.c_int32),
                ("dt", C.c_double), ("t0", C.c_double), ("mass", C.c_double), ("kT", C.c_double), ("mu", C.c_double),
                ("bandwidth", C.c_double), ("nsteps", C.c_int64), ("key", C.c_uint64), ("cell", OCell)]


class OClsDesc(C.Structure):
    _fields_ = [("save_every", C.c_int32), ("nsave", C.c_int32), ("term_kind", C.c_int32), ("pad", C.c_int32),
                ("term_x", C.c_double), ("dt", C.c_double), ("mass", C.c_double), ("nsteps", C.c_int64),
                ("cell", OCell)]


def classical_run(model, mass, r0, v0, dt=1.0, nsteps=100, save_every=0, term_kind=0, term_x=0.0, cell=None,
                  nthreads=None):
    """Simulation{Classical} trajectories by VelocityVerlet on a single-state model of one coordinate:
    final x, v, steps and per-save position, velocity, kinetic and potential energy."""
    ntraj = len(r0)
    D = OClsDesc()
    nsave = nsteps // save_every + 1 if save_every else 1
    D.save_every, D.nsave, D.term_kind, D.term_x, D.dt, D.mass, D.nsteps = (save_every, nsave, term_kind, term_x, dt,
                                                                           mass, nsteps)
    if cell is not None:
        D.cell = cell
    out = dict(x=np.ascontiguousarray(r0, dtype=np.float64).ravel().copy(),
               v=np.ascontiguousarray(v0, dtype=np.float64).ravel().copy(), steps=np.zeros(ntraj, np.int64),
               save_x=np.zeros((ntraj, nsave)), save_v=np.zeros((ntraj, nsave)), ekin=np.zeros((ntraj, nsave)),
               epot=np.zeros((ntraj, nsave)))
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    f = lib().oracle_classical_run
    f.restype = C.c_int32
    f(C.byref(D), C.byref(model), C.c_int64(ntraj), _p(out["x"]), _p(out["v"]), _p(out["steps"], C.c_int64),
      _p(out["save_x"]), _p(out["save_v"]), _p(out["ekin"]), _p(out["epot"]), C.c_int32(nthreads))
    return out


class OMapDesc(C.Structure):
    _fields_ = [("nstates", C.c_int32), ("save_every", C.c_int32), ("nsave", C.c_int32), ("method", C.c_int32),
                ("dt", C.c_double), ("mass", C.c_double), ("nsteps", C.c_int64), ("gamma", C.c_double)]


MAP_SPIN, MAP_ECMM = 0, 1


def mapping_gamma(n):
    """SpinMappingW's gamma = 2 (sqrt(n + 1) - 1) / n (spin_mapping.jl:4-9)."""
    return 2.0 / n * (np.sqrt(n + 1.0) - 1.0)


def _map_desc(n, mass, dt=0.01, nsteps=0, save_every=0, method=MAP_SPIN, gamma=0.0):
    D = OMapDesc()
    D.nstates, D.save_every, D.dt, D.mass, D.nsteps = n, save_every, dt, mass, nsteps
    D.method, D.gamma = method, gamma
    D.nsave = nsteps // save_every + 1 if save_every else 1
    return D


def mapping_sample(n, state, key, traj_ids=None):
    """Initial mapping variables of pure diabatic states (0-based `state` per trajectory): X, P (ntraj, n)."""
    st = np.ascontiguousarray(state, dtype=np.int32)
    ntraj = len(st)
    X, P = np.zeros((ntraj, n)), np.zeros((ntraj, n))
    ids = None if traj_ids is None else np.ascontiguousarray(traj_ids, dtype=np.int64)
    f = lib().oracle_mapping_sample
    f.restype = None
    f(C.c_int32(n), C.c_int64(ntraj), _p(st, C.c_int32), C.c_uint64(key), _p(ids, C.c_int64), _p(X), _p(P))
    return X, P


def ecmm_sample(n, ntraj, gamma, key, traj_ids=None):
    """eCMM's initial mapping variables: a uniform point on the 2n-sphere of radius sqrt(2 + 2 n gamma),
    positions X the first n coordinates, momenta P the last n: X, P (ntraj, n)."""
    X, P = np.zeros((ntraj, n)), np.zeros((ntraj, n))
    ids = None if traj_ids is None else np.ascontiguousarray(traj_ids, dtype=np.int64)
    f = lib().oracle_ecmm_sample
    f.restype = None
    f(C.c_int32(n), C.c_int64(ntraj), C.c_double(gamma), C.c_uint64(key), _p(ids, C.c_int64), _p(X), _p(P))
    return X, P


def mapping_energy(model, n, mass, u, method=MAP_SPIN, gamma=0.0):
    """The mapping Hamiltonian (spin mapping, or eCMM with its gamma) at u = [r, v, X(n), P(n)]."""
    D = _map_desc(n, mass, method=method, gamma=gamma)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    f = lib().oracle_mapping_energy
    f.restype = C.c_double
    return f(C.byref(D), C.byref(model), _p(uu))


def mapping_motion(model, n, mass, u, method=MAP_SPIN, gamma=0.0):
    """motion! of spin mapping (or eCMM) at u = [r, v, X(n), P(n)]: du in the same layout."""
    D = _map_desc(n, mass, method=method, gamma=gamma)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    du = np.zeros_like(uu)
    f = lib().oracle_mapping_motion
    f.restype = None
    f(C.byref(D), C.byref(model), _p(uu), _p(du))
    return du


def mapping_run(model, n, mass, r0, v0, X0, P0, dt=0.01, nsteps=100, save_every=0, nthreads=None, method=MAP_SPIN,
                gamma=0.0):
    """Spin-mapping (or eCMM) trajectories by MInt: final x, v, X, P and per save the diabatic populations
    (ntraj, nsave, n), kinetic and potential energy."""
    ntraj = len(r0)
    D = _map_desc(n, mass, dt, nsteps, save_every, method, gamma)
    out = dict(x=np.ascontiguousarray(r0, dtype=np.float64).ravel().copy(),
               v=np.ascontiguousarray(v0, dtype=np.float64).ravel().copy(),
               X=np.ascontiguousarray(X0, dtype=np.float64).reshape(ntraj, n).copy(),
               P=np.ascontiguousarray(P0, dtype=np.float64).reshape(ntraj, n).copy(),
               pop=np.zeros((ntraj, D.nsave, n)), ekin=np.zeros((ntraj, D.nsave)), epot=np.zeros((ntraj, D.nsave)),
               init=np.zeros((ntraj, n)))
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    f = lib().oracle_mapping_run
    f.restype = C.c_int32
    f(C.byref(D), C.byref(model), C.c_int64(ntraj), _p(out["x"]), _p(out["v"]), _p(out["X"]), _p(out["P"]),
      _p(out["pop"]), _p(out["ekin"]), _p(out["epot"]), _p(out["init"]), C.c_int32(nthreads))
    return out


class ORpMapDesc(C.Structure):
    _fields_ = [("nstates", C.c_int32), ("nbeads", C.c_int32), ("method", C.c_int32), ("save_every", C.c_int32),
                ("nsave", C.c_int32), ("pad", C.c_int32), ("dt", C.c_double), ("mass", C.c_double),
                ("temperature", C.c_double), ("gamma", C.c_double), ("nsteps", C.c_int64)]


RPMAP_NRPMD, RPMAP_ECMM = 0, 1


def _rpmap_desc(method, n, P, mass, temperature, gamma, dt=0.01, nsteps=0, save_every=0):
    D = ORpMapDesc()
    D.nstates, D.nbeads, D.method, D.save_every = n, P, method, save_every
    D.nsave = nsteps // save_every + 1 if save_every else 1
    D.dt, D.mass, D.temperature, D.gamma, D.nsteps = dt, mass, temperature, gamma, nsteps
    return D


def rpmap_sample(method, n, P, gamma, ntraj, key, state=None, traj_ids=None):
    """Initial mapping variables of ring polymers (NRPMD: occupied diabatic `state` per trajectory; eCMM: a point
    on the 2nP-sphere): q, p (ntraj, P, n)."""
    D = _rpmap_desc(method, n, P, 1.0, 1.0, gamma)
    q, p = np.zeros((ntraj, P, n)), np.zeros((ntraj, P, n))
    st = np.zeros(ntraj, np.int32) if state is None else np.ascontiguousarray(state, dtype=np.int32)
    ids = None if traj_ids is None else np.ascontiguousarray(traj_ids, dtype=np.int64)
    f = lib().oracle_rpmap_sample
    f.restype = None
    f(C.byref(D), C.c_int64(ntraj), _p(st, C.c_int32), C.c_uint64(key), _p(ids, C.c_int64), _p(q), _p(p))
    return q, p


def rpmap_energy(method, model, n, P, mass, temperature, gamma, u):
    """classical_hamiltonian of a mapping ring polymer at u = [r(P), v(P), q(P n), p(P n)] (bead-major)."""
    D = _rpmap_desc(method, n, P, mass, temperature, gamma)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    f = lib().oracle_rpmap_energy
    f.restype = C.c_double
    return f(C.byref(D), C.byref(model), _p(uu))


def rpmap_motion(method, model, n, P, mass, temperature, gamma, u):
    D = _rpmap_desc(method, n, P, mass, temperature, gamma)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    du = np.zeros_like(uu)
    f = lib().oracle_rpmap_motion
    f.restype = None
    f(C.byref(D), C.byref(model), _p(uu), _p(du))
    return du


def rpmap_run(method, model, n, P, mass, temperature, gamma, r0, v0, q0, p0, dt=0.01, nsteps=100, save_every=0,
              nthreads=None):
    """Ring polymers by RingPolymerMInt: final r, v (ntraj, P), q, p (ntraj, P, n) and per save the diabatic
    populations (ntraj, nsave, n), kinetic energy, potential energy with springs and centroid (ntraj, nsave)."""
    r0 = np.ascontiguousarray(r0, dtype=np.float64).reshape(-1, P)
    ntraj = r0.shape[0]
    D = _rpmap_desc(method, n, P, mass, temperature, gamma, dt, nsteps, save_every)
    out = dict(r=r0.copy(), v=np.ascontiguousarray(v0, dtype=np.float64).reshape(ntraj, P).copy(),
               q=np.ascontiguousarray(q0, dtype=np.float64).reshape(ntraj, P, n).copy(),
               p=np.ascontiguousarray(p0, dtype=np.float64).reshape(ntraj, P, n).copy(),
               pop=np.zeros((ntraj, D.nsave, n)), ekin=np.zeros((ntraj, D.nsave)), epot=np.zeros((ntraj, D.nsave)),
               centroid=np.zeros((ntraj, D.nsave)), init=np.zeros((ntraj, n)))
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    f = lib().oracle_rpmap_run
    f.restype = C.c_int32
    f(C.byref(D), C.byref(model), C.c_int64(ntraj), _p(out["r"]), _p(out["v"]), _p(out["q"]), _p(out["p"]),
      _p(out["pop"]), _p(out["ekin"]), _p(out["epot"]), _p(out["centroid"]), _p(out["init"]), C.c_int32(nthreads))
    return out


def cme_run(method, model, mass, kT, r0, v0, state0, dt=1.0, nsteps=100, mu=0.0, bandwidth=1.0, save_every=0, key=0,
            nthreads=None, cell=None):
    """CME (method 0) / BCME (1) trajectories (cme.jl); r0, v0 (ntraj,), state0 0-based diabatic states.  Returns
    final x, v, state, hops and per-save position, kinetic energy, potential V_ss and state."""
    ntraj = len(r0)
    D = OCmeDesc()
    nsave = nsteps // save_every + 1 if save_every else 1
    D.method, D.save_every, D.nsave, D.dt, D.t0, D.mass, D.kT, D.mu, D.bandwidth = (method, save_every, nsave, dt, 0.0,
                                                                                     mass, kT, mu, bandwidth)
    D.nsteps, D.key = nsteps, key
    if cell is not None:
        D.cell = cell
    out = dict(x=np.ascontiguousarray(r0, dtype=np.float64).ravel().copy(),
               v=np.ascontiguousarray(v0, dtype=np.float64).ravel().copy(),
               state=np.ascontiguousarray(state0, dtype=np.int32).copy(), nhops=np.zeros(ntraj, np.int32),
               save_x=np.zeros((ntraj, nsave)), ekin=np.zeros((ntraj, nsave)), epot=np.zeros((ntraj, nsave)),
               save_state=np.zeros((ntraj, nsave), np.int32))
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    f = lib().oracle_cme_run
    f.restype = C.c_int32
    f(C.byref(D), C.byref(model), C.c_int64(ntraj), _p(out["x"]), _p(out["v"]), _p(out["state"], C.c_int32),
      _p(out["nhops"], C.c_int32), _p(out["save_x"]), _p(out["ekin"]), _p(out["epot"]), _p(out["save_state"], C.c_int32),
      C.c_int32(nthreads))
    return out


def cme_broadening(h, mu, beta, Gam, W, force=False):
    f = lib().oracle_cme_broadening
    f.restype = C.c_double
    f.argtypes = [C.c_double] * 5 + [C.c_int32]
    return f(h, mu, beta, Gam, W, int(force))


# --- IESH / EhrenfestNA on the Newns-Anderson model ------------------------------------------
class ONADesc(C.Structure):
    _fields_ = [("mw2", C.c_double), ("g", C.c_double), ("dG", C.c_double), ("Gamma", C.c_double), ("a", C.c_double),
                ("b", C.c_double), ("M", C.c_int32), ("pad", C.c_int32)]


class OIEDesc(C.Structure):
    _fields_ = [("method", C.c_int32), ("nelec", C.c_int32), ("rescaling", C.c_int32), ("estimate", C.c_int32),
                ("dt", C.c_double), ("t0", C.c_double), ("mass", C.c_double), ("nsteps", C.c_int64), ("key", C.c_uint64),
                ("edc", C.c_int32), ("pad", C.c_int32), ("edc_C", C.c_double), ("save_every", C.c_int32), ("nsave", C.c_int32)]


def na_model(m=2000.0, omega=2e-4, g=20.6097, dG=-3.8e-3, Gamma=6.4e-3, M=40, W=None):
    """MiaoSubotnik impurity with a TrapezoidalRule bath on [-W, W] (W = 6 Gamma / 2 by default)."""
    if W is None:
        W = 6 * Gamma / 2
    d = ONADesc()
    d.mw2, d.g, d.dG, d.Gamma, d.a, d.b, d.M = m * omega ** 2, g, dG, Gamma, -W, W, M
    return d


def eigen_hermitian(H):
    H = np.asarray(H, dtype=complex)
    n = H.shape[0]
    E = np.zeros(n); Ur = np.zeros((n, n)); Ui = np.zeros((n, n))
    f = lib().oracle_eigen_hermitian
    f(C.c_int32(n), _p(np.ascontiguousarray(H.real)), _p(np.ascontiguousarray(H.imag)), _p(E), _p(Ur), _p(Ui))
    return E, Ur + 1j * Ui


def det(A):
    A = np.asarray(A, dtype=complex)
    out = np.zeros(2)
    lib().oracle_det(C.c_int32(A.shape[0]), _p(np.ascontiguousarray(A.real)), _p(np.ascontiguousarray(A.imag)), _p(out))
    return out[0] + 1j * out[1]


def na_eigen(md, x):
    N = md.M + 1
    E = np.zeros(N); U = np.zeros((N, N)); dh = C.c_double(0)
    f = lib().oracle_na_eigen
    f.argtypes = [C.POINTER(ONADesc), C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    f(C.byref(md), x, _p(E), _p(U), C.byref(dh))
    return E, U, dh.value


def fermi_dirac_occupations(E, nelec, beta, traj, key, available=None):
    """Sorted 0-based occupied states from the Boltzmann-factor chain (oracle_fermi_dirac_occupations)."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    av = np.arange(len(E), dtype=np.int32) if available is None else np.ascontiguousarray(available, dtype=np.int32)
    occ = np.zeros(nelec, np.int32)
    f = lib().oracle_fermi_dirac_occupations
    f.argtypes = [C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.c_double, C.c_uint64,
                  C.c_uint64, C.POINTER(C.c_int32)]
    f.restype = C.c_int32
    rc = f(_p(E), nelec, av.ctypes.data_as(C.POINTER(C.c_int32)), len(av), float(beta), C.c_uint64(traj),
           C.c_uint64(key), occ.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != 0:
        raise ValueError("need more available states than electrons")
    return occ


def na_matrix(md, x):
    """The Newns-Anderson model matrix at x (impurity first, then the bath levels)."""
    N = md.M + 1
    V = np.zeros((N, N))
    f = lib().oracle_na_matrix
    f.argtypes = [C.POINTER(ONADesc), C.c_double, C.POINTER(C.c_double)]
    f(C.byref(md), x, _p(V))
    return V


def fermi_dirac_diabatic(Ediab, U, nelec, beta, traj, key, available=None):
    """(sorted occupied adiabatic states, psi (N, n_e) real) of oracle_fermi_dirac_diabatic."""
    Ediab = np.ascontiguousarray(Ediab, dtype=np.float64)
    U = np.ascontiguousarray(U, dtype=np.float64)
    N = len(Ediab)
    av = np.arange(N, dtype=np.int32) if available is None else np.ascontiguousarray(available, dtype=np.int32)
    occ = np.zeros(nelec, np.int32)
    psi = np.zeros((N, nelec))
    f = lib().oracle_fermi_dirac_diabatic
    f.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                  C.c_double, C.c_uint64, C.c_uint64, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    f.restype = C.c_int32
    rc = f(_p(Ediab), _p(U), N, nelec, av.ctypes.data_as(C.POINTER(C.c_int32)), len(av), float(beta),
           C.c_uint64(traj), C.c_uint64(key), occ.ctypes.data_as(C.POINTER(C.c_int32)), _p(psi))
    if rc != 0:
        raise ValueError(f"oracle_fermi_dirac_diabatic failed ({rc})")
    return occ, psi


def iesh_run_ensemble(method, md, nelec, mass, r0, v0, occ0, psi0=None, dt=1.0, nsteps=100, key=0,
                      rescaling=RESCALE_STANDARD, estimate=True, edc=False, edc_C=0.1, save_every=0, nthreads=None):
    ntraj = len(r0)
    N = md.M + 1
    occ0 = np.ascontiguousarray(occ0, dtype=np.int32).reshape(ntraj, nelec)
    if psi0 is None:
        psi0 = np.zeros((ntraj, N, nelec), complex)
        for i in range(ntraj):
            for e in range(nelec):
                psi0[i, occ0[i, e], e] = 1.0
    R = OIEDesc()
    R.method, R.nelec, R.rescaling, R.estimate = method, nelec, rescaling, int(estimate)
    R.dt, R.t0, R.mass, R.nsteps, R.key, R.edc, R.edc_C = dt, 0.0, mass, nsteps, key, int(edc), edc_C
    nsave = nsteps // save_every + 1 if save_every else 1
    R.save_every, R.nsave = save_every, nsave
    out = dict(r=np.zeros(ntraj), v=np.zeros(ntraj), psi_re=np.zeros((ntraj, N, nelec)), psi_im=np.zeros((ntraj, N, nelec)),
               occ=np.zeros((ntraj, nelec), np.int32), nhops=np.zeros(ntraj, np.int32), pop_di=np.zeros((ntraj, nsave, N)))
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    f = lib().oracle_iesh_run_ensemble
    f.restype = C.c_int32
    f(C.byref(R), C.byref(md), C.c_int64(ntraj), _p(np.ascontiguousarray(r0, dtype=np.float64)),
      _p(np.ascontiguousarray(v0, dtype=np.float64)), _p(occ0, C.c_int32),
      _p(np.ascontiguousarray(psi0.real)), _p(np.ascontiguousarray(psi0.imag)), _p(out["r"]), _p(out["v"]),
      _p(out["psi_re"]), _p(out["psi_im"]), _p(out["occ"], C.c_int32), _p(out["nhops"], C.c_int32), _p(out["pop_di"]),
      C.c_int32(nthreads))
    out["psi"] = out["psi_re"] + 1j * out["psi_im"]
    out["pop_imp"] = out["pop_di"][:, :, 0]
    return out


def rp_iesh_run_ensemble(method, md, nelec, mass, P, temperature, r0, v0, occ0, psi0=None, dt=1.0, nsteps=100, key=0,
                         rescaling=RESCALE_STANDARD, estimate=True, save_every=0, nthreads=None):
    """Ring-polymer IESH (method 0) / EhrenfestNA (1) with BCBWavefunction; r0, v0 (ntraj, P)."""
    r0 = np.ascontiguousarray(r0, dtype=np.float64).reshape(-1, P)
    v0 = np.ascontiguousarray(v0, dtype=np.float64).reshape(-1, P)
    ntraj = r0.shape[0]
    N = md.M + 1
    occ0 = np.ascontiguousarray(occ0, dtype=np.int32).reshape(ntraj, nelec)
    if psi0 is None:
        psi0 = np.zeros((ntraj, N, nelec), complex)
        for i in range(ntraj):
            for e in range(nelec):
                psi0[i, occ0[i, e], e] = 1.0
    R = OIEDesc()
    R.method, R.nelec, R.rescaling, R.estimate = method, nelec, rescaling, int(estimate)
    R.dt, R.t0, R.mass, R.nsteps, R.key, R.edc, R.edc_C = dt, 0.0, mass, nsteps, key, 0, 0.1
    nsave = nsteps // save_every + 1 if save_every else 1
    R.save_every, R.nsave = save_every, nsave
    out = dict(r=np.zeros((ntraj, P)), v=np.zeros((ntraj, P)), psi_re=np.zeros((ntraj, N, nelec)),
               psi_im=np.zeros((ntraj, N, nelec)), occ=np.zeros((ntraj, nelec), np.int32), nhops=np.zeros(ntraj, np.int32),
               pop_di=np.zeros((ntraj, nsave, N)))
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    f = lib().oracle_rp_iesh_run_ensemble
    f.restype = C.c_int32
    f(C.byref(R), C.byref(md), C.c_int32(P), C.c_double(temperature), C.c_int64(ntraj), _p(r0), _p(v0),
      _p(occ0, C.c_int32), _p(np.ascontiguousarray(psi0.real)), _p(np.ascontiguousarray(psi0.imag)), _p(out["r"]),
      _p(out["v"]), _p(out["psi_re"]), _p(out["psi_im"]), _p(out["occ"], C.c_int32), _p(out["nhops"], C.c_int32),
      _p(out["pop_di"]), C.c_int32(nthreads))
    out["psi"] = out["psi_re"] + 1j * out["psi_im"]
    return out


def iesh_probabilities(psi, occ, d, v, dt, clamp=True):
    """prob[m, n] for electron n hopping to unoccupied state m (no early exit)."""
    psi = np.asarray(psi, dtype=complex)
    N, Ne = psi.shape
    out = np.zeros((N, Ne))
    lib().oracle_iesh_probabilities(C.c_int32(N), C.c_int32(Ne), _p(np.ascontiguousarray(psi.real)),
                                    _p(np.ascontiguousarray(psi.imag)), _p(np.ascontiguousarray(occ, dtype=np.int32), C.c_int32),
                                    _p(np.ascontiguousarray(d, dtype=np.float64)), C.c_double(v), C.c_double(dt),
                                    C.c_int32(int(clamp)), _p(out))
    return out
