"""ctypes binding of libnqcd_mi355x.so (the C ABI in include/nqcd_mi355x.h).

This is the product path: there is no CPU fallback.  Loading fails loudly when
the shared library has not been built.
"""
import ctypes as C
import os

import numpy as np

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("NQCD_MI355X_LIB", os.path.join(_PKG_ROOT, "libnqcd_mi355x.so"))

ABI_VERSION = 2

# enum nqcd_status
OK, ERR_ARG, ERR_UNSUPPORTED, ERR_DEVICE, ERR_STATE, ERR_NOMEM = 0, -1, -2, -3, -4, -5
# enum nqcd_method
(METHOD_FSSH, METHOD_EHRENFEST, METHOD_IESH, METHOD_EHRENFEST_NA, METHOD_RPMD, METHOD_RPSH, METHOD_RP_EHRENFEST,
 METHOD_CME, METHOD_BCME, METHOD_CLASSICAL, METHOD_SPIN_MAPPING, METHOD_ECMM, METHOD_NRPMD) = range(13)
# enum nqcd_algorithm
(ALG_BAB_TSIT5, ALG_VERLET_ELECTRONICS, ALG_BCB, ALG_BCB_TSIT5, ALG_BCB_WAVEFUNCTION, ALG_TSIT5,
 ALG_VELOCITY_VERLET, ALG_MINT, ALG_RP_MINT) = range(9)
RESCALE_STANDARD, RESCALE_VINVERSION, RESCALE_OFF = range(3)
PREC_FP64, PREC_FP32_NUCLEI = range(2)
SCHED_PERSISTENT, SCHED_STEP_PER_LAUNCH = range(2)
(MODEL_TULLY1, MODEL_TULLY2, MODEL_TULLY3, MODEL_SPIN_BOSON, MODEL_DW_GAUSSIAN, MODEL_HARMONIC, MODEL_NEWNS_ANDERSON,
 MODEL_DIATOM_SURFACE, MODEL_MLP_SURROGATE, MODEL_MORSE_N, MODEL_ANDERSON_HOLSTEIN_1) = range(11)
MLP_PARAMS = 128 * 6 + 128 + 128 * 128 + 128 + 3 * 128 + 3
DECOHERENCE_NONE, DECOHERENCE_EDC = range(2)
TERM_NONE, TERM_ABS_GT, TERM_GT, TERM_LT = range(4)
CELL_NONE, CELL_BOUNDARY, CELL_TERMINATE = range(3)
(SAMPLER_GAUSSIAN, SAMPLER_TULLY_GRID, SAMPLER_HARMONIC_WIGNER, SAMPLER_RING_POLYMER_THERMAL,
 SAMPLER_RING_POLYMER_PIMC) = range(5)
(OUT_ADIABATIC_POPULATION, OUT_DIABATIC_POPULATION, OUT_SCATTERING_1D, OUT_POPULATION_CORRELATION, OUT_FINAL_ADIABATIC_POPULATION,
 OUT_CENTROID_POSITION, OUT_KINETIC_ENERGY, OUT_TOTAL_ENERGY, OUT_SCATTERING_1D_DIABATIC, OUT_POSITION,
 OUT_VELOCITY) = range(11)
RED_MEAN, RED_SUM = range(2)
(FIELD_R, FIELD_V, FIELD_SIGMA_RE, FIELD_SIGMA_IM, FIELD_STATE, FIELD_STEPS, FIELD_NHOPS, FIELD_SAVE_POP_AD,
 FIELD_SAVE_POP_DI, FIELD_SAVE_STATE, FIELD_SAVE_CENTROID, FIELD_FALLBACKS, FIELD_SAVE_R, FIELD_SAVE_EKIN,
 FIELD_SAVE_EPOT, FIELD_FINAL_POP_DI, FIELD_SAVE_V, FIELD_MAP_X, FIELD_MAP_P, FIELD_INIT_POP_DI) = range(20)

EXPORTED_SYMBOLS = (
    "nqcd_mi355x_abi_version", "nqcd_mi355x_open", "nqcd_mi355x_close", "nqcd_mi355x_set_problem",
    "nqcd_mi355x_set_model", "nqcd_mi355x_load_initial", "nqcd_mi355x_sample_initial", "nqcd_mi355x_run",
    "nqcd_mi355x_reduce", "nqcd_mi355x_download", "nqcd_mi355x_last_run_stats", "nqcd_mi355x_last_error",
    "nqcd_mi355x_surrogate_train", "nqcd_mi355x_surrogate_eval", "nqcd_mi355x_surrogate_gradient",
    "nqcd_mi355x_surrogate_set_params", "nqcd_mi355x_surrogate_get_params", "nqcd_mi355x_test_determinant",
    "nqcd_mi355x_sample_fermi_dirac", "nqcd_mi355x_autotune", "nqcd_mi355x_reduce_noalloc",
    "nqcd_mi355x_test_broadening", "nqcd_mi355x_set_cell", "nqcd_mi355x_load_mapping",
)


class ProblemDesc(C.Structure):
    _fields_ = [("method", C.c_int32), ("algorithm", C.c_int32), ("n_states", C.c_int32), ("ndofs", C.c_int32),
                ("natoms", C.c_int32), ("nbeads", C.c_int32), ("n_electrons", C.c_int32), ("rescaling", C.c_int32),
                ("precision", C.c_int32), ("schedule", C.c_int32), ("electronic_substeps", C.c_int32),
                ("decoherence", C.c_int32), ("masses", C.POINTER(C.c_double)), ("dt", C.c_double),
                ("temperature", C.c_double), ("edc_C", C.c_double), ("disable_estimate", C.c_int32),
                ("reserved0", C.c_int32), ("bandwidth", C.c_double), ("fermi_level", C.c_double),
                ("mapping_gamma", C.c_double)]


class TermDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("dof", C.c_int32), ("value", C.c_double)]


class SamplerDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("initial_state", C.c_int32), ("state_basis", C.c_int32), ("n_models", C.c_int32),
                ("ntraj", C.c_int64), ("params", C.POINTER(C.c_double)), ("nparams", C.c_int64)]


class TrainDesc(C.Structure):
    _fields_ = [("steps", C.c_int32), ("batch", C.c_int32), ("lr", C.c_double), ("lr_final", C.c_double),
                ("loss_weights", C.c_double * 3), ("seed", C.c_uint64)]


class RunStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("wall_ms", C.c_double), ("traj_steps", C.c_int64), ("launches", C.c_int64),
                ("block", C.c_int32), ("grid", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("block", C.c_int32), ("workgroups_per_cu", C.c_int32), ("n_candidates", C.c_int32),
                ("family", C.c_int32), ("best_ms", C.c_double), ("default_ms", C.c_double)]


class NQCDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[nqcd_mi355x status {code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Load libnqcd_mi355x.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libnqcd_mi355x.so not found at {LIB_PATH}: build it with "
                          f"`python nqcdynamics.jl_amd/build.py` (the MI355X backend has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    H = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    L.nqcd_mi355x_abi_version.restype = C.c_int32
    L.nqcd_mi355x_open.argtypes = [C.c_int32, C.POINTER(H)]
    L.nqcd_mi355x_close.argtypes = [H]
    L.nqcd_mi355x_close.restype = None
    L.nqcd_mi355x_set_problem.argtypes = [H, C.POINTER(ProblemDesc)]
    L.nqcd_mi355x_set_model.argtypes = [H, C.c_int32, C.c_int32, dp, C.c_int64]
    L.nqcd_mi355x_set_cell.argtypes = [H, C.c_int32, C.c_int32, dp, ip]
    L.nqcd_mi355x_load_initial.argtypes = [H, C.c_int64, dp, dp, dp, dp, ip, ip]
    L.nqcd_mi355x_load_mapping.argtypes = [H, C.c_int64, dp, dp]
    L.nqcd_mi355x_sample_initial.argtypes = [H, C.POINTER(SamplerDesc), C.c_uint64]
    L.nqcd_mi355x_sample_fermi_dirac.argtypes = [H, C.c_double, ip, C.c_int32, C.c_int32, C.c_uint64]
    L.nqcd_mi355x_run.argtypes = [H, C.c_double, C.c_double, C.c_int32, C.POINTER(TermDesc), C.c_uint64]
    L.nqcd_mi355x_reduce.argtypes = [H, C.c_int32, C.c_int32, dp, C.c_int64]
    L.nqcd_mi355x_reduce_noalloc.argtypes = [H, C.c_int32, C.c_int32, dp, C.c_int64]
    L.nqcd_mi355x_download.argtypes = [H, C.c_int32, C.c_void_p, C.c_int64]
    L.nqcd_mi355x_last_run_stats.argtypes = [H, C.POINTER(RunStats)]
    L.nqcd_mi355x_autotune.argtypes = [H, C.c_double, C.c_double, C.POINTER(TermDesc), C.c_uint64, C.POINTER(TuneResult)]
    fp = C.POINTER(C.c_float)
    L.nqcd_mi355x_surrogate_train.argtypes = [H, dp, dp, C.c_int64, C.POINTER(TrainDesc), dp]
    L.nqcd_mi355x_surrogate_eval.argtypes = [H, C.c_int32, dp, C.c_int64, dp, dp]
    L.nqcd_mi355x_surrogate_gradient.argtypes = [H, C.c_int32, dp, dp, C.c_int64, dp, dp, dp]
    L.nqcd_mi355x_surrogate_set_params.argtypes = [H, C.c_int32, fp, fp]
    L.nqcd_mi355x_surrogate_get_params.argtypes = [H, C.c_int32, fp, fp]
    L.nqcd_mi355x_test_determinant.argtypes = [H, dp, dp, C.c_int32, C.c_int32, dp]
    L.nqcd_mi355x_test_broadening.argtypes = [H, dp, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp]
    L.nqcd_mi355x_last_error.argtypes = [H]
    L.nqcd_mi355x_last_error.restype = C.c_char_p
    for name in EXPORTED_SYMBOLS:
        if name not in ("nqcd_mi355x_close", "nqcd_mi355x_last_error", "nqcd_mi355x_abi_version"):
            getattr(L, name).restype = C.c_int32
    v = L.nqcd_mi355x_abi_version()
    if v != ABI_VERSION:
        raise ImportError(f"libnqcd_mi355x ABI version {v} != expected {ABI_VERSION}")
    _lib = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))


class Ensemble:
    """One device ensemble: a thin owner of an nqcd_handle_t."""

    def __init__(self, device=0):
        self._L = lib()
        self._h = C.c_void_p()
        self._check(self._L.nqcd_mi355x_open(device, C.byref(self._h)), None)
        self.ntraj = 0
        self.n_states = 0
        self.D = 0
        self.rows = 0
        self.nbeads = 1
        self.nsave = 1

    def _check(self, rc, h="self"):
        if rc != OK:
            handle = self._h if h == "self" else None
            msg = self._L.nqcd_mi355x_last_error(handle)
            raise NQCDError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.nqcd_mi355x_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_problem(self, method, n_states, ndofs, natoms, masses, dt, algorithm=ALG_BAB_TSIT5, nbeads=1,
                    n_electrons=0, rescaling=RESCALE_STANDARD, precision=PREC_FP64, schedule=SCHED_PERSISTENT,
                    electronic_substeps=5, temperature=0.0, decoherence=DECOHERENCE_NONE, edc_C=0.1,
                    estimate_probability=True, bandwidth=0.0, fermi_level=0.0, mapping_gamma=0.0):
        self._masses = np.ascontiguousarray(np.broadcast_to(np.asarray(masses, dtype=np.float64), (natoms,)))
        d = ProblemDesc()
        d.method, d.algorithm, d.n_states, d.ndofs, d.natoms, d.nbeads = method, algorithm, n_states, ndofs, natoms, nbeads
        d.n_electrons, d.rescaling, d.precision, d.schedule = n_electrons, rescaling, precision, schedule
        d.electronic_substeps, d.masses, d.dt, d.temperature = electronic_substeps, _dp(self._masses), dt, temperature
        d.decoherence, d.edc_C, d.disable_estimate = decoherence, edc_C, 0 if estimate_probability else 1
        d.bandwidth, d.fermi_level, d.mapping_gamma = bandwidth, fermi_level, mapping_gamma
        self._check(self._L.nqcd_mi355x_set_problem(self._h, C.byref(d)))
        self.method, self.n_electrons = method, n_electrons
        self.n_states, self.D, self.dt = n_states, ndofs * natoms, dt
        self.nbeads = nbeads
        self.rows = ndofs * natoms * (nbeads if method in (METHOD_RPMD, METHOD_RPSH, METHOD_RP_EHRENFEST) or
                                      algorithm == ALG_RP_MINT else 1)
        self.map_beads = nbeads if algorithm == ALG_RP_MINT else 0   # 0: no bead axis

    def set_model(self, slot, kind, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._check(self._L.nqcd_mi355x_set_model(self._h, slot, kind, _dp(p), len(p)))

    def set_cell(self, mode, vectors=None, periodic=None):
        """Simulation cell of the cell callbacks: mode CELL_BOUNDARY (wrap) or CELL_TERMINATE, vectors the
        ndofs x ndofs cell matrix with the lattice vectors as columns, periodic ndofs flags (None: all);
        CELL_NONE clears it."""
        if mode == CELL_NONE:
            self._check(self._L.nqcd_mi355x_set_cell(self._h, CELL_NONE, 0, None, None))
            return
        a = np.asarray(vectors, dtype=np.float64)
        a = a.reshape(1, 1) if a.ndim == 0 else a
        nd = a.shape[0]
        col = np.ascontiguousarray(a.T)   # column-major
        per = None if periodic is None else np.ascontiguousarray(np.asarray(periodic, dtype=np.int32).ravel())
        self._check(self._L.nqcd_mi355x_set_cell(self._h, mode, nd, _dp(col), _ip(per)))

    def load_initial(self, r, v, sig, state=None, model_id=None):
        r = np.ascontiguousarray(r, dtype=np.float64)
        ntraj = r.shape[0]
        v = np.ascontiguousarray(v, dtype=np.float64)
        sr = si = None
        if sig is not None:  # None for problems without a density (RPMD, CME / BCME)
            sig = np.asarray(sig)
            # Julia column-major (n, n, ntraj) == C row-major (ntraj, n, n) transposed per matrix
            sr = np.ascontiguousarray(np.real(sig).transpose(0, 2, 1), dtype=np.float64)
            si = np.ascontiguousarray(np.imag(sig).transpose(0, 2, 1), dtype=np.float64)
        st = None if state is None else np.ascontiguousarray(state, dtype=np.int32)
        mid = None if model_id is None else np.ascontiguousarray(model_id, dtype=np.int32)
        self._check(self._L.nqcd_mi355x_load_initial(self._h, ntraj, _dp(r), _dp(v), _dp(sr), _dp(si), _ip(st), _ip(mid)))
        self.ntraj = ntraj

    def load_initial_iesh(self, r, v, occ, psi=None):
        """IESH / EhrenfestNA: r, v (ntraj,) or, for ring polymers, (ntraj, nbeads); occ (ntraj,
        n_electrons) occupied adiabatic states (0-based), psi (ntraj, n_states, n_electrons) complex or
        None for psi[occ_e, e] = 1."""
        r = np.asarray(r, dtype=np.float64)
        ntraj = r.shape[0]
        r = np.ascontiguousarray(np.ravel(r))
        v = np.ascontiguousarray(np.ravel(np.asarray(v, dtype=np.float64)))
        occ = np.ascontiguousarray(occ, dtype=np.int32).reshape(ntraj, -1)
        pr = pi = None
        if psi is not None:
            psi = np.asarray(psi)
            # Julia column-major (n, ne, ntraj) == C row-major (ntraj, ne, n)
            pr = np.ascontiguousarray(np.real(psi).transpose(0, 2, 1), dtype=np.float64)
            pi = np.ascontiguousarray(np.imag(psi).transpose(0, 2, 1), dtype=np.float64)
        self._check(self._L.nqcd_mi355x_load_initial(self._h, ntraj, _dp(r), _dp(v), _dp(pr), _dp(pi), _ip(occ), None))
        self.ntraj = ntraj

    def load_mapping(self, X, P):
        """Spin mapping / eCMM: mapping positions X and momenta P (ntraj, n), or (ntraj, nbeads, n) for ring
        polymers, of the loaded trajectories (None for both: draw them in the run from the run's key)."""
        if X is None and P is None:
            self._check(self._L.nqcd_mi355x_load_mapping(self._h, self.ntraj, None, None))
            return
        X = np.ascontiguousarray(X, dtype=np.float64)
        P = np.ascontiguousarray(P, dtype=np.float64)
        self._check(self._L.nqcd_mi355x_load_mapping(self._h, X.shape[0], _dp(X), _dp(P)))

    def sample_initial(self, kind, ntraj, params, seed, initial_state=0, state_basis=0, n_models=1):
        p = np.ascontiguousarray(params, dtype=np.float64)
        s = SamplerDesc()
        s.kind, s.initial_state, s.state_basis, s.n_models, s.ntraj = kind, initial_state, state_basis, n_models, ntraj
        s.params, s.nparams = _dp(p), len(p)
        self._check(self._L.nqcd_mi355x_sample_initial(self._h, C.byref(s), C.c_uint64(seed)))
        self.ntraj = ntraj

    def sample_fermi_dirac(self, beta, seed, available_states=None, diabatic=False):
        """IESH / EhrenfestNA: occupations of the loaded trajectories from the Fermi-Dirac chain at
        inverse temperature beta (numpy.inf for 0 K); available_states are 0-based (None = all);
        diabatic=True samples diabatic states and builds the orbitals from them (IESH)."""
        av = None if available_states is None else np.ascontiguousarray(available_states, dtype=np.int32)
        self._check(self._L.nqcd_mi355x_sample_fermi_dirac(self._h, float(beta), _ip(av), 0 if av is None else len(av),
                                                           1 if diabatic else 0, C.c_uint64(seed)))

    @staticmethod
    def _term(term):
        """TermDesc pointer of (dof, value[, kind]) (kind TERM_ABS_GT by default), or None."""
        if term is None:
            return None
        td = TermDesc()
        td.kind, td.dof, td.value = (term[2] if len(term) > 2 else TERM_ABS_GT), term[0], term[1]
        return C.pointer(td)

    def run(self, t0, t1, key, save_every=0, term=None):
        self._check(self._L.nqcd_mi355x_run(self._h, t0, t1, save_every, self._term(term), C.c_uint64(key)))
        nsteps = int(round((t1 - t0) / self.dt))
        self.nsave = nsteps // save_every + 1 if save_every > 0 else 1
        return self.stats()

    def autotune(self, t0, t1, key, term=None):
        """Time the persistent kernel's candidate launch shapes on [t0, t1] of the loaded problem (with the
        termination predicate `term`, as for run) and keep the fastest for later runs
        (nqcd_mi355x_autotune); the handle's results are overwritten."""
        r = TuneResult()
        self._check(self._L.nqcd_mi355x_autotune(self._h, t0, t1, self._term(term), C.c_uint64(key), C.byref(r)))
        return {k: getattr(r, k) for k, _ in TuneResult._fields_}

    def stats(self):
        s = RunStats()
        self._check(self._L.nqcd_mi355x_last_run_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in RunStats._fields_}

    def reduce(self, output_kind, reduction=RED_MEAN, length=None):
        out = np.zeros(length, dtype=np.float64)
        self._check(self._L.nqcd_mi355x_reduce(self._h, output_kind, reduction, _dp(out), length))
        return out

    def reduce_noalloc(self, output_kind, reduction=RED_MEAN, length=None):
        """The same reduction through the entry point that allocates no device memory."""
        out = np.zeros(length, dtype=np.float64)
        self._check(self._L.nqcd_mi355x_reduce_noalloc(self._h, output_kind, reduction, _dp(out), length))
        return out

    def download(self, field):
        if getattr(self, "method", None) in (METHOD_IESH, METHOD_EHRENFEST_NA):
            return self._download_iesh(field)
        n, D, nt, nsave = max(self.n_states, 2), self.rows, self.ntraj, self.nsave
        shapes = {FIELD_R: ((nt, D), np.float64), FIELD_V: ((nt, D), np.float64),
                  FIELD_SAVE_CENTROID: ((nt, nsave), np.float64),
                  FIELD_SIGMA_RE: ((nt, n, n), np.float64), FIELD_SIGMA_IM: ((nt, n, n), np.float64),
                  FIELD_STATE: ((nt,), np.int32), FIELD_STEPS: ((nt,), np.int32), FIELD_NHOPS: ((nt,), np.int32),
                  FIELD_SAVE_POP_AD: ((nt, nsave, n), np.float64), FIELD_SAVE_POP_DI: ((nt, nsave, n), np.float64),
                  FIELD_SAVE_STATE: ((nt, nsave), np.int32), FIELD_FALLBACKS: ((nt,), np.int32),
                  FIELD_SAVE_R: ((nt, nsave, D), np.float64), FIELD_SAVE_EKIN: ((nt, nsave), np.float64),
                  FIELD_SAVE_EPOT: ((nt, nsave), np.float64), FIELD_FINAL_POP_DI: ((nt, n), np.float64),
                  FIELD_SAVE_V: ((nt, nsave), np.float64)}
        shapes[FIELD_INIT_POP_DI] = ((nt, n), np.float64)
        mb = getattr(self, "map_beads", 0)
        shapes[FIELD_MAP_X] = shapes[FIELD_MAP_P] = (((nt, n) if mb == 0 else (nt, mb, n)), np.float64)
        if getattr(self, "method", None) in (METHOD_CME, METHOD_BCME, METHOD_CLASSICAL):  # one coordinate
            shapes[FIELD_SAVE_R] = ((nt, nsave), np.float64)
        shape, dtype = shapes[field]
        out = np.zeros(shape, dtype=dtype)
        self._check(self._L.nqcd_mi355x_download(self._h, field, out.ctypes.data_as(C.c_void_p), out.nbytes))
        if field in (FIELD_SIGMA_RE, FIELD_SIGMA_IM):
            out = out.transpose(0, 2, 1).copy()  # column-major (n, n) per trajectory -> row-major
        return out

    # --- MLP surrogate ----------------------------------------------------------------------------
    def surrogate_train(self, X, Y, steps=2000, batch=1024, lr=3e-3, lr_final=1e-4, loss_weights=(1.0, 1.0, 1.0),
                        seed=0x4E5143444D493335):
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        d = TrainDesc()
        d.steps, d.batch, d.lr, d.lr_final, d.seed = steps, batch, lr, lr_final, seed
        for i in range(3):
            d.loss_weights[i] = loss_weights[i]
        hist = np.zeros(2 * steps)
        self._check(self._L.nqcd_mi355x_surrogate_train(self._h, _dp(X), _dp(Y), len(X), C.byref(d), _dp(hist)))
        return hist.reshape(2, steps)

    def surrogate_eval(self, member, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        Hout = np.zeros((len(X), 3)); dH = np.zeros((len(X), 3, 6))
        self._check(self._L.nqcd_mi355x_surrogate_eval(self._h, member, _dp(X), len(X), _dp(Hout), _dp(dH)))
        return Hout, dH

    def surrogate_gradient(self, member, X, Y, w=(1.0, 1.0, 1.0)):
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        wv = np.ascontiguousarray(w, dtype=np.float64)
        g = np.zeros(MLP_PARAMS); loss = np.zeros(1)
        self._check(self._L.nqcd_mi355x_surrogate_gradient(self._h, member, _dp(X), _dp(Y), len(X), _dp(wv), _dp(g), _dp(loss)))
        return g, float(loss[0])

    def surrogate_set_params(self, member, params, norm=None):
        p = np.ascontiguousarray(params, dtype=np.float32)
        nv = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        fp = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))
        self._check(self._L.nqcd_mi355x_surrogate_set_params(self._h, member, fp(p), fp(nv)))

    def surrogate_get_params(self, member):
        p = np.zeros(MLP_PARAMS, np.float32); nv = np.zeros(18, np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        self._check(self._L.nqcd_mi355x_surrogate_get_params(self._h, member, fp(p), fp(nv)))
        return p, nv

    def test_broadening(self, hs, mu, kT, Gamma, W):
        """The device's BCME broadening integrals n(h), n2(h) (test hook)."""
        hs = np.ascontiguousarray(hs, dtype=np.float64)
        n, n2 = np.zeros_like(hs), np.zeros_like(hs)
        self._check(self._L.nqcd_mi355x_test_broadening(self._h, _dp(hs), len(hs), mu, kT, Gamma, W, _dp(n), _dp(n2)))
        return n, n2

    def test_determinant(self, S):
        """Determinants of complex matrices S (nmat, n, n) by the device routine of the IESH
        hopping estimate -> (det (nmat,) complex, ok (nmat,) bool)."""
        S = np.asarray(S, dtype=np.complex128)
        nmat, n = S.shape[0], S.shape[1]
        re = np.ascontiguousarray(S.real); im = np.ascontiguousarray(S.imag)
        out = np.zeros((nmat, 3))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self._check(self._L.nqcd_mi355x_test_determinant(self._h, dp(re), dp(im), n, nmat, dp(out)))
        return out[:, 0] + 1j * out[:, 1], out[:, 2] != 0

    def _download_iesh(self, field):
        n, ne, nt, nsave = self.n_states, self.n_electrons, self.ntraj, self.nsave
        rv = (nt,) if self.nbeads == 1 else (nt, self.nbeads)
        shapes = {FIELD_R: (rv, np.float64), FIELD_V: (rv, np.float64),
                  FIELD_SIGMA_RE: ((nt, ne, n), np.float64), FIELD_SIGMA_IM: ((nt, ne, n), np.float64),
                  FIELD_STATE: ((nt, ne), np.int32), FIELD_STEPS: ((nt,), np.int32), FIELD_NHOPS: ((nt,), np.int32),
                  FIELD_SAVE_POP_AD: ((nt, nsave, n), np.float64), FIELD_SAVE_POP_DI: ((nt, nsave, n), np.float64)}
        shape, dtype = shapes[field]
        out = np.zeros(shape, dtype=dtype)
        self._check(self._L.nqcd_mi355x_download(self._h, field, out.ctypes.data_as(C.c_void_p), out.nbytes))
        if field in (FIELD_SIGMA_RE, FIELD_SIGMA_IM):
            out = out.transpose(0, 2, 1).copy()  # (ntraj, n, ne)
        return out
