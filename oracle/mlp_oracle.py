"""CPU oracle of the troubled-cell selector network (TEST INFRASTRUCTURE ONLY).

FP64 numpy forward/backward of the 75 -> 64 -> 64 -> 64 -> 1 GELU/sigmoid network with the
same parameter layout as the library (FDMAMR_SELECTOR_NPARAM floats). With `bf16=True` the
operands are rounded to BF16 exactly where the GPU kernels feed matrix cores (weights,
hidden activations of layers 1-2, back-propagated deltas), so the library's gradients can be
compared with it; with `bf16=False` it is the exact FP64 network, pinned by finite
differences. The network has no reference counterpart.
"""
from __future__ import annotations

import numpy as np

IN, H = 80, 64
P_W1 = 0
P_B1 = P_W1 + IN * H
P_W2 = P_B1 + H
P_B2 = P_W2 + H * H
P_W3 = P_B2 + H
P_B3 = P_W3 + H * H
P_W4 = P_B3 + H
P_B4 = P_W4 + H
P_TOTAL = P_B4 + 1


def bf16_round(x):
    """float -> nearest-even BF16 -> float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits_to_float(b):
    return (np.asarray(b, np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(w):
    w = np.asarray(w, np.float64)
    return dict(W1=w[P_W1:P_B1].reshape(IN, H), b1=w[P_B1:P_W2], W2=w[P_W2:P_B2].reshape(H, H),
                b2=w[P_B2:P_W3], W3=w[P_W3:P_B3].reshape(H, H), b3=w[P_B3:P_W4], w4=w[P_W4:P_B4],
                b4=w[P_B4])


# The activation the library evaluates (csrc/fdmamr_selector.hip, gelu / gelu_d): GELU as
# x s(c), c = x clamped to [-GELU_A, GELU_A], s(c) = 1/2 + c P(c^2), P the degree-5 minimax fit
# of the tanh form's factor 1 / (1 + exp(-2 k0 (x + k1 x^3))) (k0 = sqrt(2/pi), k1 = 0.044715)
# with s(GELU_A) = 1; the same FP32 coefficients, evaluated in FP64.
GELU_K0 = float(np.sqrt(2.0 / np.pi))
GELU_K1 = 0.044715
GELU_A = 3.5
GELU_P = [float.fromhex(h) for h in ("0x1.958016p-2", "-0x1.fed952p-5", "0x1.013ab0p-7",
                                     "-0x1.496018p-11", "0x1.da140ap-16", "-0x1.215ff2p-21")]


def gelu_tanh(x):
    """GELU in its tanh form, 0.5 x (1 + tanh(k0 (x + k1 x^3)))."""
    x = np.asarray(x, np.float64)
    return 0.5 * x * (1.0 + np.tanh(GELU_K0 * (x + GELU_K1 * x ** 3)))


def _poly(t, c):
    p = np.zeros_like(t) + c[-1]
    for a in c[-2::-1]:
        p = p * t + a
    return p


def gelu(x):
    x = np.asarray(x, np.float64)
    c = np.clip(x, -GELU_A, GELU_A)
    return x * (0.5 + c * _poly(c * c, GELU_P))


def gelu_d(x):
    x = np.asarray(x, np.float64)
    c = np.clip(x, -GELU_A, GELU_A)
    t = c * c
    s = 0.5 + c * _poly(t, GELU_P)
    dp = _poly(t, [k * a for k, a in enumerate(GELU_P)][1:])
    ds = _poly(t, GELU_P) + 2.0 * t * dp
    return np.where(np.abs(x) <= GELU_A, s + x * ds, s)


def forward(w, X, bf16=True):
    """X: (n, 80) float features. Returns probabilities and the cache."""
    p = unpack(w)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    x = r(X)
    z1 = x @ r(p["W1"]) + p["b1"]
    h1 = r(gelu(z1))
    z2 = h1 @ r(p["W2"]) + p["b2"]
    h2 = r(gelu(z2))
    z3 = h2 @ r(p["W3"]) + p["b3"]
    h3 = gelu(z3)
    z4 = h3 @ p["w4"] + p["b4"]
    prob = 1.0 / (1.0 + np.exp(-z4))
    return prob, dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, z4=z4, p=p, r=r)


def loss_and_grad(w, X, y, pos_weight=1.0, bf16=True):
    prob, c = forward(w, X, bf16)
    r, p = c["r"], c["p"]
    y = np.asarray(y, np.float64)
    n = len(y)
    pc = np.clip(prob, 1e-7, 1 - 1e-7)
    loss = float(np.mean(-(pos_weight * y * np.log(pc) + (1 - y) * np.log(1 - pc))))
    dz4 = (prob * (pos_weight * y + 1 - y) - pos_weight * y) / n
    g = np.zeros(P_TOTAL)
    g[P_W4:P_B4] = c["h3"].T @ dz4
    g[P_B4] = dz4.sum()
    dz3 = np.outer(dz4, p["w4"]) * gelu_d(c["z3"])
    g[P_B3:P_W4] = dz3.sum(0)
    dz3b = r(dz3)
    g[P_W3:P_B3] = (c["h2"].T @ dz3b).reshape(-1)
    dz2 = (dz3b @ r(p["W3"]).T) * gelu_d(c["z2"])
    g[P_B2:P_W3] = dz2.sum(0)
    dz2b = r(dz2)
    g[P_W2:P_B2] = (c["h1"].T @ dz2b).reshape(-1)
    dz1 = (dz2b @ r(p["W2"]).T) * gelu_d(c["z1"])
    g[P_B1:P_W2] = dz1.sum(0)
    gw1 = c["x"].T @ r(dz1)
    gw1[75:] = 0.0
    g[P_W1:P_B1] = gw1.reshape(-1)
    return loss, g


def random_weights(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    w = np.zeros(P_TOTAL)
    w[P_W1:P_B1] = (rng.normal(size=(IN, H)) * np.sqrt(2 / 75) * scale).reshape(-1)
    w[P_W1 + 75 * H:P_B1] = 0.0
    w[P_B1:P_W2] = rng.normal(size=H) * 0.1
    w[P_W2:P_B2] = rng.normal(size=H * H) * np.sqrt(2 / H) * scale
    w[P_B2:P_W3] = rng.normal(size=H) * 0.1
    w[P_W3:P_B3] = rng.normal(size=H * H) * np.sqrt(2 / H) * scale
    w[P_B3:P_W4] = rng.normal(size=H) * 0.1
    w[P_W4:P_B4] = rng.normal(size=H) * np.sqrt(1 / H)
    w[P_B4] = 0.1
    return w
