"""nqcd_mi355x_reduce_noalloc (the reduction entry point that allocates no device memory) against
nqcd_mi355x_reduce on the same device results: bit-identical for every output kind, on two runs in a
row (the second smaller, so the shared scratch array holds the first call's values beyond what the
second uses), and a clear error past the scratch limit."""
import numpy as np
import pytest

from nqcd_mi355x import _lib as L

pytestmark = pytest.mark.gpu

KEY = 0x4E5143444D493335
TULLY1 = [0.01, 1.6, 0.005, 1.0]


def _run(E, method, ntraj, nsteps, save_every, seed):
    mass = 2000.0
    rng = np.random.default_rng(seed)
    r0 = rng.normal(-4.0, 0.6, size=(ntraj, 1))
    v0 = (rng.uniform(8, 30, size=ntraj) / mass)[:, None]
    sig0 = np.zeros((ntraj, 2, 2), complex); sig0[:, 0, 0] = 1
    E.set_problem(method, 2, 1, 1, [mass], 1.0)
    E.set_model(0, L.MODEL_TULLY1, TULLY1)
    E.load_initial(r0, v0, sig0, state=np.zeros(ntraj, np.int32))
    E.run(0.0, float(nsteps), KEY, save_every=save_every)


def _kinds(E):
    n, ns = 2, E.nsave
    return [(L.OUT_ADIABATIC_POPULATION, n * ns), (L.OUT_DIABATIC_POPULATION, n * ns),
            (L.OUT_FINAL_ADIABATIC_POPULATION, n), (L.OUT_SCATTERING_1D, 2 * n),
            (L.OUT_POPULATION_CORRELATION, n * n * ns), (L.OUT_KINETIC_ENERGY, ns), (L.OUT_TOTAL_ENERGY, ns)]


def _check_all(E, label):
    for kind, length in _kinds(E):
        for red in (L.RED_MEAN, L.RED_SUM):
            print(f"{label}: output {kind} reduction {red} ({length} values)", flush=True)
            a = E.reduce(kind, red, length)
            b = E.reduce_noalloc(kind, red, length)
            assert np.array_equal(a, b), (label, kind, red, np.abs(a - b).max())
            assert np.all(np.isfinite(b))


@pytest.mark.parametrize("method", [L.METHOD_FSSH, L.METHOD_EHRENFEST])
def test_reduce_noalloc_matches_reduce_two_calls(method):
    with L.Ensemble(0) as E:
        # 40000 trajectories: 10 chunks of 4096 per row; 201 save slots
        _run(E, method, 40000, 400, 2, seed=1)
        _check_all(E, "first")
        # the second call is smaller (fewer trajectories, chunks and rows)
        _run(E, method, 3000, 60, 6, seed=2)
        _check_all(E, "second")


def test_reduce_noalloc_rejects_beyond_scratch():
    with L.Ensemble(0) as E:
        # 2^17 trajectories -> 32 chunks; 2049 slots -> 8196 correlation rows: 262272 partials > 262144
        _run(E, L.METHOD_FSSH, 1 << 17, 2048, 1, seed=3)
        n4 = 4 * E.nsave
        with pytest.raises(L.NQCDError) as e:
            E.reduce_noalloc(L.OUT_POPULATION_CORRELATION, L.RED_MEAN, n4)
        assert e.value.code == L.ERR_ARG and "scratch" in str(e.value)
        # within the limit on the same results: 4098 population rows x 32 chunks
        a = E.reduce(L.OUT_ADIABATIC_POPULATION, L.RED_MEAN, 2 * E.nsave)
        b = E.reduce_noalloc(L.OUT_ADIABATIC_POPULATION, L.RED_MEAN, 2 * E.nsave)
        assert np.array_equal(a, b)
        assert np.allclose(a.reshape(-1, 2).sum(axis=1), 1.0, atol=1e-12)
