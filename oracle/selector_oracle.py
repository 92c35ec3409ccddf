"""CPU oracle of the troubled-cell selector's inputs (TEST INFRASTRUCTURE ONLY).

The selector has no reference counterpart; its label and features are defined on top of the
reference's numerics:

* label (per interior leaf): the largest deviation |w1 - w1_lin| of the first WCNS_Z3 weight
  from its linear value d1 / (d1 + d2) over every reconstruction the leaf's flux derivative
  makes: 2 directions x 2 faces x 4 characteristic fields x 4 reconstructions (w and g, left
  and right, nnd_center_RK.F90:87-104, 208-225), with the weights of nnd_center_RK.F90:326-346
  except that the smoothness indicators are floored at 1e-6 of the three values' mean square
  (the reference's epsilon is 1e-40, so on flat data its weights follow round-off) and the
  ratio tau3 / (floor + IS) is capped at 1e30;
* features (per interior leaf, 75 = 25 points x 3 channels): the leaf's 5x5 stencil at its own
  spacing, the cross (+-1, +-2 in x and y) being the solver's own stencil values
  (XuanGua.F90:401-590), the 16 off-axis points the leaf containing the point; channels rho,
  p and |u|; each channel centred on the leaf's value c and divided by max(m, |c|) + 1e-3 |c|,
  m the largest |v - c| over the stencil; feature index 3 * ((b + 2) * 5 + (a + 2)) + channel
  for the point at (a dx, b dy).

Both are written in FP64 numpy from these definitions; the library computes them in FP32.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

D1 = float(np.float32(1.0) / np.float32(11.0))    # REAL(4) 1./11.  nnd_center_RK.F90:331
D2 = float(np.float32(10.0) / np.float32(11.0))   # REAL(4) 10./11.
W1_LIN = D1 / (D1 + D2)
FLOOR_REL = 1e-6
RATIO_CAP = 1e30


def weight_deviation(u1, u2, u3):
    """|w1 - w1_lin| of WCNS_Z3 (nnd_center_RK.F90:326-346) with the floored indicators;
    arrays of any shape."""
    u1, u2, u3 = (np.asarray(a, np.float64) for a in (u1, u2, u3))
    IS1 = np.minimum((-u1 + u2) ** 2, (-u2 + u3) ** 2)
    IS2 = (u1 - 2.0 * u2 + u3) ** 2 + u1 * u1 + u2 * u2 + u3 * u3 - u1 * u2 - u2 * u3 - u1 * u3
    tau3 = (10.0 * np.abs(IS2 - 3.0 * IS1)) ** 2
    fl = FLOOR_REL * (u1 * u1 + u2 * u2 + u3 * u3) / 3.0 + 1e-30
    a1 = D1 * (1.0 + np.minimum(tau3 / (fl + IS1), RATIO_CAP))
    a2 = D2 * (1.0 + np.minimum(tau3 / (fl + IS2), RATIO_CAP))
    return np.abs(a1 / (a1 + a2) - W1_LIN)


def _cons_flux(q, gamma, direction):
    """compute_E / compute_F (nnd_center_RK.F90:171-185,293-307): q[..., (rho,u,v,T,p)]."""
    r, u, v, p = q[..., 0], q[..., 1], q[..., 2], q[..., 4]
    E = p / (gamma - 1.0) + 0.5 * r * (u * u + v * v)
    U = np.stack([r, r * u, r * v, E], -1)
    vn = u if direction == 0 else v
    F = np.stack([r * vn, r * u * vn + (p if direction == 0 else 0.0),
                  r * v * vn + (p if direction == 1 else 0.0), vn * (E + p)], -1)
    return U, F


def labels(center, stencil, gamma):
    """center: (N, 5) primitive state; stencil: (N, 8, 5) the +-1, +-2 values in the library's
    order (+1x, -1x, +1y, -1y, +2x, -2x, +2y, -2y). Returns (N,) labels."""
    N = len(center)
    out = np.zeros(N)
    cv = 1.0 / (gamma * (gamma - 1.0))
    cols = {0: [5, 1, None, 0, 4], 1: [7, 3, None, 2, 6]}  # i-2, i-1, i, i+1, i+2
    for direction in (0, 1):
        pts = np.stack([center if k is None else stencil[:, k] for k in cols[direction]], 1)  # (N, 5, 5)
        U, F = _cons_flux(pts, gamma, direction)
        for f in range(2):  # face i-1/2 (points 0..3), i+1/2 (points 1..4)
            avg = (U[:, f + 1] + U[:, f + 2]) / 2.0
            Ri = np.empty((N, 4, 4))
            for n in range(N):
                Ri[n] = O.eigen(direction, avg[n], gamma, cv)[1]
            w = np.einsum("nrc,njc->njr", Ri, U[:, f:f + 4])
            g = np.einsum("nrc,njc->njr", Ri, F[:, f:f + 4])
            for a in (w, g):
                dev = np.maximum(weight_deviation(a[:, 0], a[:, 1], a[:, 2]),
                                 weight_deviation(a[:, 3], a[:, 2], a[:, 1]))
                out = np.maximum(out, dev.max(axis=1))
    return out


def channels(prim):
    """rho, p, |u| of primitive states (..., 5)."""
    return np.stack([prim[..., 0], prim[..., 4], np.sqrt(prim[..., 1] ** 2 + prim[..., 2] ** 2)], -1)


def features(center, points25):
    """center: (N, 5); points25: (N, 25, 5) states of the 5x5 grid, point (b+2)*5 + (a+2)
    (the centre at 12). Returns (N, 75) features."""
    c = channels(center)[:, None, :]
    v = channels(points25)
    m = np.max(np.abs(v - c), axis=1, keepdims=True)
    scale = np.maximum(m, np.abs(c)) + 1e-3 * np.abs(c) + 1e-30
    return ((v - c) / scale).reshape(len(center), 75)
