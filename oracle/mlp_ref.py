"""CPU reference of the MLP surrogate of the diabatic Hamiltonian (test infrastructure only).

Network 6 -> 128 -> 128 -> 3 with tanh activations on normalised inputs:
    xn = (x - in_shift) * in_scale
    a1 = W1 xn + b1,  h1 = tanh(a1)
    a2 = W2 h1 + b2,  h2 = tanh(a2)
    o  = W3 h2 + b3
    (H11, H22, H12) = o * out_scale + out_shift
Precision points of the device kernels, reproduced when `bf16=True`: the first layer in FP32,
h1 and h2 and the W2, W3 weights rounded to BF16 before their products (FP32 accumulation), the
tangents of the forward-mode input gradient rounded to BF16 where they enter the W2 / W3 products.
Loss: mean over samples of sum_o w_o (o - t_o)^2 on normalised targets t = (y - out_shift) / out_scale.
The network's inputs are the descriptor q(x) of the configuration x = (x1, y1, z1, x2, y2, z2):
    q = (b, z1, z2, (z1 - z2) / b, (x1 + x2) / 2, (y1 + y2) / 2),  b = |r1 - r2|,
normalised as xn = (q - in_shift) * in_scale; forces follow by the chain rule through dq/dx.
"""
import numpy as np

H = 128


def bf16(a):
    """Round-to-nearest-even to bfloat16, returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounding = ((u >> 16) & 1) + 0x7FFF
    u = ((u + rounding) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def descriptor(X):
    """(n, 6) configurations -> (n, 6) descriptors, FP64."""
    X = np.atleast_2d(np.asarray(X, dtype=np.float64))
    d = X[:, :3] - X[:, 3:]
    b = np.linalg.norm(d, axis=1)
    return np.stack([b, X[:, 2], X[:, 5], d[:, 2] / b, 0.5 * (X[:, 0] + X[:, 3]), 0.5 * (X[:, 1] + X[:, 4])], axis=1)


def descriptor_jacobian(X):
    """dq/dx, shape (n, 6, 6), FP64, written out from the definition entry by entry."""
    X = np.atleast_2d(np.asarray(X, dtype=np.float64))
    n = len(X)
    d = X[:, :3] - X[:, 3:]
    b = np.linalg.norm(d, axis=1)
    D = np.zeros((n, 6, 6))
    for a in range(3):
        D[:, 0, a] = d[:, a] / b
        D[:, 0, 3 + a] = -d[:, a] / b
        dc = ((a == 2) - d[:, 2] * d[:, a] / b ** 2) / b
        D[:, 3, a] = dc
        D[:, 3, 3 + a] = -dc
    D[:, 1, 2] = 1.0
    D[:, 2, 5] = 1.0
    D[:, 4, 0] = D[:, 4, 3] = 0.5
    D[:, 5, 1] = D[:, 5, 4] = 0.5
    return D


def init(seed):
    rng = np.random.default_rng(seed)
    P = {"W1": rng.normal(0, 1 / np.sqrt(6), (H, 6)), "b1": np.zeros(H),
         "W2": rng.normal(0, 1 / np.sqrt(H), (H, H)), "b2": np.zeros(H),
         "W3": rng.normal(0, 1 / np.sqrt(H), (3, H)), "b3": np.zeros(3)}
    return {k: v.astype(np.float32) for k, v in P.items()}


def forward(P, xn, bf=False):
    """xn: (n, 6) normalised inputs -> dict with a1, h1, a2, h2, o (all (n, .))."""
    xn = np.asarray(xn, dtype=np.float64)
    W2 = bf16(P["W2"]).astype(np.float64) if bf else P["W2"].astype(np.float64)
    W3 = bf16(P["W3"]).astype(np.float64) if bf else P["W3"].astype(np.float64)
    a1 = xn @ P["W1"].astype(np.float64).T + P["b1"]
    h1 = np.tanh(a1)
    h1q = bf16(h1).astype(np.float64) if bf else h1
    a2 = h1q @ W2.T + P["b2"]
    h2 = np.tanh(a2)
    h2q = bf16(h2).astype(np.float64) if bf else h2
    o = h2q @ W3.T + P["b3"]
    return {"a1": a1, "h1": h1, "h1q": h1q, "a2": a2, "h2": h2, "h2q": h2q, "o": o, "W2": W2, "W3": W3}


def jacobian(P, xn, bf=False):
    """d o / d xn, shape (n, 3, 6), in reverse mode (the device's formulation): per output o,
    delta2 = W3[o] * (1 - h2^2), delta1 = (W2^T delta2) * (1 - h1^2), d o / d xn = W1^T delta1.
    With bf16=True, delta2 and delta1 are rounded to BF16 before their products and W3, W2, W1 are the
    BF16 weights (the device runs both products on BF16 MFMA)."""
    f = forward(P, xn, bf)
    s1 = 1.0 - f["h1"] ** 2
    s2 = 1.0 - f["h2"] ** 2
    W1 = P["W1"].astype(np.float64)
    J = np.zeros((len(f["o"]), 3, 6))
    for o in range(3):
        d2 = s2 * f["W3"][o]
        if bf:
            d2 = bf16(d2).astype(np.float64)
        d1 = (d2 @ f["W2"]) * s1
        if bf:
            d1 = bf16(d1).astype(np.float64)
        J[:, o, :] = d1 @ (bf16(P["W1"]).astype(np.float64) if bf else W1)
    return J


def loss_and_grads(P, xn, t, w=(1.0, 1.0, 1.0)):
    """FP64 loss and exact parameter gradients (no BF16 rounding)."""
    w = np.asarray(w, dtype=np.float64)
    n = len(xn)
    f = forward(P, xn)
    r = f["o"] - t
    loss = np.mean(np.sum(w * r * r, axis=1))
    d3 = 2.0 * w * r / n
    g = {"W3": d3.T @ f["h2"], "b3": d3.sum(0)}
    d2 = (d3 @ P["W3"].astype(np.float64)) * (1 - f["h2"] ** 2)
    g["W2"] = d2.T @ f["h1"]
    g["b2"] = d2.sum(0)
    d1 = (d2 @ P["W2"].astype(np.float64)) * (1 - f["h1"] ** 2)
    g["W1"] = d1.T @ np.asarray(xn, dtype=np.float64)
    g["b1"] = d1.sum(0)
    return loss, g
