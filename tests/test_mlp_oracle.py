"""Pins for the selector network's CPU oracle (oracle/mlp_oracle.py).

The FP64 network (bf16=False) is checked against central finite differences of its own loss,
which fixes every backward formula (a transposed weight, a missing activation derivative or a
dropped bias term fails it). The BF16-rounding variant must agree with it to BF16 precision.
"""
import numpy as np

from oracle import mlp_oracle as M


def _data(n=37, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, size=(n, M.IN))
    X[:, 75:] = 0.0
    y = (rng.uniform(size=n) < 0.3).astype(np.uint8)
    return X, y


def test_fp64_gradient_matches_finite_differences():
    X, y = _data()
    w = M.random_weights(3)
    loss, g = M.loss_and_grad(w, X, y, pos_weight=3.0, bf16=False)
    rng = np.random.default_rng(7)
    idx = np.concatenate([rng.choice(np.arange(M.P_W1, M.P_W1 + 75 * M.H), 12, replace=False),
                          rng.choice(np.arange(M.P_B1, M.P_TOTAL), 40, replace=False), [M.P_B4]])
    eps = 1e-6
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += eps
        wm[k] -= eps
        lp, _ = M.loss_and_grad(wp, X, y, 3.0, bf16=False)
        lm, _ = M.loss_and_grad(wm, X, y, 3.0, bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, g[k])
    # padded input rows carry no gradient
    assert np.all(g[M.P_W1 + 75 * M.H:M.P_B1] == 0.0)


def test_bf16_variant_close_to_fp64():
    X, y = _data(200)
    w = M.random_weights(4)
    l64, g64 = M.loss_and_grad(w, X, y, 2.0, bf16=False)
    l16, g16 = M.loss_and_grad(w, X, y, 2.0, bf16=True)
    assert abs(l64 - l16) < 2e-2 * abs(l64)
    assert np.linalg.norm(g16 - g64) < 5e-2 * np.linalg.norm(g64)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159], np.float32)
    r = M.bf16_round(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0          # tie -> even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12


def test_gelu_closed_forms():
    from scipy.special import erf
    x = np.linspace(-8, 8, 4001)
    g = M.gelu(x)
    # the polynomial form stays within 6.2e-4 of the tanh form and within 1e-3 of the exact
    # GELU x Phi(x)
    assert np.max(np.abs(g - M.gelu_tanh(x))) < 6.2e-4
    exact = 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))
    assert np.max(np.abs(g - exact)) < 1e-3
    assert M.gelu(0.0) == 0.0 and M.gelu_d(0.0) == 0.5
    # odd part: g(x) - g(-x) = x
    assert np.allclose(g - M.gelu(-x), x, atol=1e-14)
    # tails: identity for large x, zero for large -x (s(GELU_A) = 1 up to the FP32 coefficients)
    assert abs(M.gelu(30.0) - 30.0) < 1e-5 and abs(M.gelu(-30.0)) < 1e-5
    assert M.gelu_d(30.0) == M.gelu_d(M.GELU_A + 1e-9) and abs(M.gelu_d(30.0) - 1.0) < 1e-6
    # the sigmoid factor is monotone between 0 and 1
    s = M.gelu(x[x > 0]) / x[x > 0]
    assert np.all(np.diff(s) >= -1e-15) and np.all((s >= 0.5) & (s <= 1.0 + 1e-6))
    # the derivative by central differences (away from the clamp points, where g' jumps)
    h = 1e-6
    xs = x[np.abs(np.abs(x) - M.GELU_A) > 1e-3]
    fd = (M.gelu(xs + h) - M.gelu(xs - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_d(xs))) < 1e-8
