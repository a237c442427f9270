"""The MLP surrogate on the GPU: BF16 forward / input-gradient kernels and the training gradient
against the CPU reference (oracle/mlp_ref.py), training on configurations visited by FSSH, and
FSSH driven by the trained surrogate against FSSH on the analytic model."""
import os
import sys

import numpy as np
import pytest

import nqcd_mi355x as N

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402

pytestmark = pytest.mark.gpu
IDENT = np.concatenate([np.zeros(6), np.ones(6), np.zeros(3), np.ones(3)]).astype(np.float32)


def _flat(P):
    return np.concatenate([P["W1"].ravel(), P["b1"], P["W2"].ravel(), P["b2"], P["W3"].ravel(), P["b3"]]).astype(np.float32)


def _random_params(seed):
    rng = np.random.default_rng(seed)
    P = mlp_ref.init(seed)
    P["b1"] = rng.normal(0, 0.2, 128).astype(np.float32)
    P["b2"] = rng.normal(0, 0.2, 128).astype(np.float32)
    P["b3"] = rng.normal(0, 0.2, 3).astype(np.float32)
    return P


def _configs(n, seed):
    """Diatom configurations: bond 2.2 +- 0.3 bohr, random orientation, height 1.5 .. 8 bohr."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3)); u /= np.linalg.norm(u, axis=1, keepdims=True)
    b = 2.2 + 0.3 * rng.normal(size=(n, 1))
    c = np.concatenate([rng.normal(size=(n, 2)), rng.uniform(1.5, 8.0, size=(n, 1))], axis=1)
    return np.concatenate([c + 0.5 * b * u, c - 0.5 * b * u], axis=1)


def _norm_of(X):
    q = mlp_ref.descriptor(X)
    return np.concatenate([q.mean(0), 1.0 / q.std(0), np.zeros(3), np.ones(3)]).astype(np.float32)


def test_forward_and_input_gradient_match_cpu_reference():
    from nqcd_mi355x import _lib as L
    P = _random_params(11)
    X = _configs(1000, 12)
    norm = _norm_of(X)
    E = L.Ensemble(0)
    E.surrogate_set_params(0, _flat(P), norm)
    Hg, Jg = E.surrogate_eval(0, X)
    E.close()
    # the device rounds X to FP32 before the descriptor
    Xf = X.astype(np.float32).astype(np.float64)
    xn = (mlp_ref.descriptor(Xf) - norm[:6]) * norm[6:12]
    Dq = mlp_ref.descriptor_jacobian(Xf) * norm[6:12][None, :, None]   # d xn / d x
    ref = mlp_ref.forward(P, xn, bf=True)["o"]
    Jref = mlp_ref.jacobian(P, xn, bf=True) @ Dq
    exact, Jexact = mlp_ref.forward(P, xn)["o"], mlp_ref.jacobian(P, xn) @ Dq
    # same rounding points: differences are FP32 tanh / accumulation order flipping a BF16 rounding
    assert np.abs(Hg - ref).max() < 5e-3 and np.sqrt(np.mean((Hg - ref) ** 2)) < 5e-4
    assert np.abs(Jg - Jref).max() < 2e-2 * np.abs(Jref).max()
    assert np.sqrt(np.mean((Jg - Jref) ** 2)) < 2e-3 * np.sqrt(np.mean(Jref ** 2))
    # and the BF16 network stays close to the exact one
    assert np.sqrt(np.mean((Hg - exact) ** 2)) < 1e-2 * np.sqrt(np.mean(exact ** 2))
    assert np.sqrt(np.mean((Jg - Jexact) ** 2)) < 2e-2 * np.sqrt(np.mean(Jexact ** 2))


def test_training_gradient_matches_cpu_reference():
    from nqcd_mi355x import _lib as L
    P = _random_params(21)
    rng = np.random.default_rng(22)
    X = _configs(512, 23); Y = rng.normal(size=(512, 3))
    norm = _norm_of(X)
    w = (1.0, 0.5, 2.0)
    E = L.Ensemble(0)
    E.surrogate_set_params(1, _flat(P), norm)
    g, loss = E.surrogate_gradient(1, X, Y, w)
    E.close()
    Pd = {k: v.astype(np.float64) for k, v in P.items()}
    xn = (mlp_ref.descriptor(X) - norm[:6]) * norm[6:12]
    lref, gref = mlp_ref.loss_and_grads(Pd, xn, Y, w)
    gref = _flat({k: v for k, v in gref.items()}).astype(np.float64)
    assert abs(loss - lref) < 2e-2 * lref
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    assert rel < 3e-2, rel
    for a, b in ((0, 768), (768, 896), (896, 17280), (17280, 17408), (17408, 17792), (17792, 17795)):
        cos = g[a:b] @ gref[a:b] / (np.linalg.norm(g[a:b]) * np.linalg.norm(gref[a:b]))
        assert cos > 0.995, (a, b, cos)


def test_coincident_atoms_rejected():
    from nqcd_mi355x import _lib as L
    E = L.Ensemble(0)
    E.surrogate_set_params(0, _flat(_random_params(3)), IDENT)
    with pytest.raises(RuntimeError, match="coincident"):
        E.surrogate_eval(0, np.zeros((4, 6)))
    E.close()


def _inputs(ntraj, seed, Z=6.0, vz=3e-3):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(ntraj, 3)); u /= np.linalg.norm(u, axis=1, keepdims=True)
    b = 2.2 + 0.1 * rng.normal(size=(ntraj, 1))
    c = np.concatenate([rng.normal(size=(ntraj, 2)), np.full((ntraj, 1), Z)], axis=1)
    r0 = np.concatenate([c + 0.5 * b * u, c - 0.5 * b * u], axis=1)
    v0 = 2e-4 * rng.normal(size=(ntraj, 6))
    v0[:, 2] -= vz; v0[:, 5] -= vz
    return r0, v0


def _fssh(L, model, r0, v0, key, precision=1, save_every=0, nsteps=5000):
    ntraj = len(r0)
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_FSSH, 2, 3, 2, [12859.0, 12859.0], 2.0, precision=precision)
    E.set_model(0, model.kind, model.params())
    if model.kind == L.MODEL_MLP_SURROGATE:
        for k, p in enumerate(model.members):
            E.surrogate_set_params(k, p, model.norm)
    sig = np.zeros((ntraj, 2, 2), complex); sig[:, 0, 0] = 1
    E.load_initial(r0, v0, sig, state=np.zeros(ntraj, np.int32))
    st = E.run(0.0, 2.0 * nsteps, key, save_every=save_every, term=(2, 8.0))
    out = {"final": E.reduce(L.OUT_FINAL_ADIABATIC_POPULATION, L.RED_MEAN, 2), "stats": st}
    if save_every:
        out["configs"] = E.download(L.FIELD_SAVE_R)
        out["steps"] = E.download(L.FIELD_STEPS)
    if model.kind == L.MODEL_MLP_SURROGATE:
        out["fallbacks"] = E.download(L.FIELD_FALLBACKS)
    E.close()
    return out


def test_surrogate_trained_on_fssh_configurations_reproduces_populations(key):
    """Train on the configurations plain FSSH visits, then run FSSH on the surrogate: the final
    population is within 0.01 of the analytic-model run on the same initial conditions and hop
    uniforms, and training is what makes it so (an untrained network is far off)."""
    from nqcd_mi355x import _lib as L
    analytic = N.DiatomSurface()
    # configurations visited by the plain (analytic) method
    r0, v0 = _inputs(4000, 41)
    vis = _fssh(L, analytic, r0, v0, key, save_every=20)
    nsave = vis["configs"].shape[1]
    keep = np.arange(nsave)[None, :] * 20 <= vis["steps"][:, None]
    X = vis["configs"][keep]
    assert len(X) > 200000
    model, hist = N.train_surrogate(analytic, X, steps=40000, batch=2048, lr=3e-3, lr_final=1e-5, threshold=1e-3)
    print(f"training loss: first {hist[:, :10].mean():.4g}  last {hist[:, -100:].mean():.4g}")
    assert hist[:, -100:].mean() < 2e-3 * hist[:, :10].mean()
    # held-out accuracy of the ensemble mean
    Xt = X[np.random.default_rng(1).choice(len(X), 5000, replace=False)] + 0.01 * np.random.default_rng(2).normal(size=(5000, 6))
    E = L.Ensemble(0)
    for k, p in enumerate(model.members):
        E.surrogate_set_params(k, p, model.norm)
    Hm = 0.5 * (E.surrogate_eval(0, Xt)[0] + E.surrogate_eval(1, Xt)[0])
    E.close()
    err = np.sqrt(np.mean((Hm - analytic.potential(Xt)) ** 2, axis=0))
    print("held-out RMS error of (H11, H22, H12):", err)
    assert np.all(err < 1e-3), err
    # FSSH on new initial conditions
    r1, v1 = _inputs(20000, 43)
    ref = _fssh(L, analytic, r1, v1, key)
    sur = _fssh(L, model, r1, v1, key)
    p_ref, p_sur = ref["final"][1], sur["final"][1]
    sigma = np.sqrt(p_ref * (1 - p_ref) / len(r1))
    fb = sur["fallbacks"].sum() / sur["stats"]["traj_steps"]
    print(f"final excited population: analytic {p_ref:.4f}  surrogate {p_sur:.4f}  sigma {sigma:.4f}  fallback fraction {fb:.4f}")
    assert abs(p_sur - p_ref) < 0.01
    assert fb < 0.5
    # an untrained pair of networks (random initial weights) is far from the analytic populations
    raw = N.MLPSurrogate(analytic, [_flat(mlp_ref.init(5)), _flat(mlp_ref.init(6))], model.norm, threshold=1e9)
    bad = _fssh(L, raw, r1[:4000], v1[:4000], key)
    assert not abs(bad["final"][1] - p_ref) <= 0.05
