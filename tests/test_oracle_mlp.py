"""The CPU reference of the MLP surrogate: parameter gradients and the forward-mode input gradient
against central finite differences, and the BF16 rounding helper against its definition."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 3.0e-5], dtype=np.float32)
    r = mlp_ref.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # 1 + 2^-8 is half an ulp: tie to even
    assert r[3] == 1.0                      # tie to even (mantissa ...0)
    assert r[2] == np.float32(1.0 + 2 ** -7)  # 1 + 1.5 ulp rounds up to 1 + 2 ulp
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-5) / 3.0e-5 < 2 ** -8


def test_parameter_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    P = mlp_ref.init(1)
    P = {k: v.astype(np.float64) for k, v in P.items()}
    P["b1"] = rng.normal(0, 0.1, 128); P["b2"] = rng.normal(0, 0.1, 128); P["b3"] = rng.normal(0, 0.1, 3)
    X = rng.normal(size=(20, 6)); T = rng.normal(size=(20, 3))
    w = (1.0, 0.5, 2.0)
    loss, g = mlp_ref.loss_and_grads(P, X, T, w)
    h = 1e-6
    for name, idxs in (("W1", [(0, 0), (77, 5)]), ("b1", [(3,)]), ("W2", [(5, 9), (127, 0)]), ("b2", [(64,)]),
                       ("W3", [(2, 100)]), ("b3", [(1,)])):
        for ix in idxs:
            Pp = {k: v.copy() for k, v in P.items()}; Pm = {k: v.copy() for k, v in P.items()}
            Pp[name][ix] += h; Pm[name][ix] -= h
            fd = (mlp_ref.loss_and_grads(Pp, X, T, w)[0] - mlp_ref.loss_and_grads(Pm, X, T, w)[0]) / (2 * h)
            assert abs(fd - g[name][ix]) < 1e-6 * max(1.0, abs(fd)), (name, ix, fd, g[name][ix])


def test_reverse_mode_input_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    P = {k: v.astype(np.float64) for k, v in mlp_ref.init(3).items()}
    X = rng.normal(size=(5, 6))
    J = mlp_ref.jacobian(P, X)
    h = 1e-6
    for i in range(6):
        Xp, Xm = X.copy(), X.copy(); Xp[:, i] += h; Xm[:, i] -= h
        fd = (mlp_ref.forward(P, Xp)["o"] - mlp_ref.forward(P, Xm)["o"]) / (2 * h)
        assert np.allclose(J[:, :, i], fd, atol=1e-7)
    # the BF16 rounding points move the gradient by the BF16 resolution only
    Jb = mlp_ref.jacobian({k: v.astype(np.float32) for k, v in P.items()}, X, bf=True)
    assert np.abs(Jb - J).max() < 3e-2 * np.abs(J).max()


def test_descriptor_jacobian_matches_finite_differences_and_invariances():
    rng = np.random.default_rng(7)
    X = rng.normal(size=(50, 6)) + np.array([0, 0, 5.0, 0, 0, 3.0])
    D = mlp_ref.descriptor_jacobian(X)
    h = 1e-6
    for i in range(6):
        e = np.zeros(6); e[i] = h
        fd = (mlp_ref.descriptor(X + e) - mlp_ref.descriptor(X - e)) / (2 * h)
        assert np.abs(fd - D[:, :, i]).max() < 1e-8
    # rotation about the surface normal and translation parallel to it leave (b, z1, z2, cos) unchanged
    c, s = np.cos(0.7), np.sin(0.7)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    Xr = np.concatenate([X[:, :3] @ R.T, X[:, 3:] @ R.T], axis=1) + np.array([1.5, -2.0, 0, 1.5, -2.0, 0])
    q, qr = mlp_ref.descriptor(X), mlp_ref.descriptor(Xr)
    assert np.allclose(q[:, :4], qr[:, :4], atol=1e-12)
    # known values
    q1 = mlp_ref.descriptor([[0, 0, 3.0, 0, 4.0, 0]])[0]
    assert np.allclose(q1, [5.0, 3.0, 0.0, 0.6, 0.0, 2.0])
