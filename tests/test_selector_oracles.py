"""The selector's inputs: label and 5x5 features (oracle/selector_oracle.py), pinned on the CPU by
closed forms, then compared with the library's FP32 kernels on the GPU (fdmamr_labels,
fdmamr_selector_features) on a mesh with level jumps.

Tolerances: the library evaluates in FP32 from an FP32 copy of the state (WLS-WENO level-jump
values included), the oracle in FP64. Features (delivered as BF16) agree to two BF16 units in
the last place plus 1e-5. Labels: the weight deviation is ill-conditioned at some leaves (its
ratios tau3 / IS carry the units of the characteristic variables squared; relative input
changes of 1e-7 move some labels by 0.2), so each leaf's tolerance is 2e-3 plus twice the
spread of the oracle's label under relative input perturbations of 3e-7 (FP32 rounding and a
few ulps of arithmetic), on the library's own FP32 stencil inputs; on the oracle's FP64 stencil
values 99% of the labels agree to 1e-2.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import selector_oracle as S
from oracle import mlp_oracle as M
import fdmamr as F
from helpers import oracle_kwargs_from_config, riemann_c2
from test_oracle_stencil_pins import Model


# ---------------------------------------------------------------- CPU pins
def _reference_w1(u1, u2, u3, eps=float(np.float32(1e-40))):
    """w1 of WCNS_Z3 with the reference's own epsilon (nnd_center_RK.F90:319-344)."""
    IS1 = min((-u1 + u2) ** 2, (-u2 + u3) ** 2)
    IS2 = (u1 - 2 * u2 + u3) ** 2 + u1 * u1 + u2 * u2 + u3 * u3 - u1 * u2 - u2 * u3 - u1 * u3
    tau3 = (10 * abs(IS2 - 3 * IS1)) ** 2
    a1 = S.D1 * (1 + tau3 / (eps + IS1))
    a2 = S.D2 * (1 + tau3 / (eps + IS2))
    return a1 / (a1 + a2)


def test_weight_deviation_closed_forms():
    # constant and linear data: tau3 = 0, the weights are the linear ones
    assert S.weight_deviation(1.5, 1.5, 1.5) == 0.0
    assert S.weight_deviation(1.0, 1.25, 1.5) == 0.0
    # a step after u2: IS1 = 0, IS2 = 2 -> w1 -> 1, the first-order candidate
    assert S.weight_deviation(0.0, 0.0, 1.0) == pytest.approx(1.0 - S.W1_LIN, abs=1e-5)
    # resolved data (IS far above the floor): the reference's own deviation
    for u in [(0.1, 0.4, 0.2), (1.0, 2.0, 5.0), (-0.3, 0.9, 0.85)]:
        assert S.weight_deviation(*u) == pytest.approx(abs(_reference_w1(*u) - S.W1_LIN), rel=1e-5)
    # flat up to round-off: the reference's weights follow the 1e-16 noise, the floor does not
    u = (1.3, 1.3 + 2e-16, 1.3)
    assert abs(_reference_w1(*u) - S.W1_LIN) > 0.5
    assert S.weight_deviation(*u) < 1e-12


def test_labels_zero_on_uniform_and_linear_states():
    N = 7
    c = np.tile([1.2, 0.3, -0.2, 0.0, 0.9], (N, 1))
    st = np.tile(c[:, None, :], (1, 8, 1))
    assert np.all(S.labels(c, st, 1.4) == 0.0)
    # linear density along x, constant velocity and pressure: every characteristic field is
    # linear along each column, so every weight is linear
    d = np.array([1.0, -1.0, 0.0, 0.0, 2.0, -2.0, 0.0, 0.0])
    st = np.tile(c[:, None, :], (1, 8, 1))
    st[:, :, 0] = 1.2 + 0.01 * d[None, :]
    assert np.all(S.labels(c, st, 1.4) < 1e-9)
    # a density jump two cells to the right is seen
    st[:, 4, 0] = 3.0
    assert np.all(S.labels(c, st, 1.4) > 0.5)


def test_features_closed_form_on_a_linear_field():
    # rho = 1 + 0.2 x - 0.1 y, p = 2, |u| = 0.5 on the 5x5 grid of spacing 0.1 about (1, 1):
    # rho - c = 0.1 (0.2 a - 0.1 b); the largest is 0.1 * 0.6 = 0.06 < c = 1.1, so the scale is
    # c + 1e-3 c; p and |u| are uniform (features 0)
    a, b = np.meshgrid(np.arange(-2, 3), np.arange(-2, 3))
    a, b = a.reshape(-1), b.reshape(-1)
    pts = np.zeros((1, 25, 5))
    pts[0, :, 0] = 1.0 + 0.2 * (1 + 0.1 * a) - 0.1 * (1 + 0.1 * b)
    pts[0, :, 1] = 0.3
    pts[0, :, 2] = 0.4
    pts[0, :, 4] = 2.0
    c = pts[:, 12]
    f = S.features(c, pts).reshape(25, 3)
    want = 0.1 * (0.2 * a - 0.1 * b) / (1.1 * 1.001)
    assert np.allclose(f[:, 0], want, rtol=1e-12, atol=1e-15)
    assert np.all(f[:, 1:] == 0.0)
    assert np.abs(f).max() <= 1.0


# ---------------------------------------------------------------- GPU parity
def _mesh():
    cfg = riemann_c2(h=1.0 / 64, lvl_amr=2, pre_adapt=2)
    o = O.Oracle(**oracle_kwargs_from_config(cfg))
    for _ in range(3):
        assert o.step() == 0
    return cfg, o


def _product(o, cfg):
    s = F.Solver(riemann_c2(h=1.0 / 64, lvl_amr=2, pre_adapt=2, flux_precision=F.FLUX_FP32))
    m, n = o.dims()
    s.import_tree(m, n, o.tree())
    s.build_tables()
    return s


def _oracle_inputs(o, cfg):
    L = o.leaves()
    st = o.stencil_all(values=True)
    center = L["prim"]
    cross = st["var"][:, :, :5]
    return L, st, center, cross


@pytest.mark.gpu
def test_labels_match_oracle():
    cfg, o = _mesh()
    s = _product(o, cfg)
    L, st, center, cross = _oracle_inputs(o, cfg)
    got = s.labels().astype(np.float64)
    # the library's own FP32 stencil inputs (level-jump values and ghosts as its stage reads them)
    cross32 = np.zeros_like(cross)
    for i in range(len(center)):
        v = s.stencil_values_fast(i).astype(np.float64)  # (8, 4): rho, u, v, p
        cross32[i, :, [0, 1, 2, 4]] = v.T
    c32 = center.astype(np.float32).astype(np.float64)
    want32 = S.labels(c32, cross32, cfg.gamma)
    rng = np.random.default_rng(0)
    spread = np.zeros(len(got))
    for _ in range(6):
        cp = c32 * (1 + 3e-7 * rng.standard_normal(c32.shape))
        xp = cross32 * (1 + 3e-7 * rng.standard_normal(cross32.shape))
        spread = np.maximum(spread, np.abs(S.labels(cp, xp, cfg.gamma) - want32))
    tol = 2e-3 + 2.0 * spread
    err32 = np.abs(got - want32)
    want = S.labels(center, cross, cfg.gamma)
    err = np.abs(got - want)
    print("labels: leaves", len(got), "troubled", int((want > 0.05).sum()), "on the library's inputs: max err",
          err32.max(), "max err / tol", float((err32 / tol).max()), "leaves with spread > 1e-2",
          int((spread > 1e-2).sum()), "on the FP64 stencil: 99% quantile", np.quantile(err, 0.99))
    assert (want > 0.05).sum() > 100
    assert np.all(err32 <= tol)
    assert np.quantile(err, 0.99) < 1e-2


@pytest.mark.gpu
def test_features_match_oracle():
    cfg, o = _mesh()
    s = _product(o, cfg)
    L, st, center, cross = _oracle_inputs(o, cfg)
    model = Model(o, cfg.h, cfg.xmin, cfg.ymin)
    t = model.t
    N = len(center)
    pts = np.zeros((N, 25, 5))
    pts[:, 12] = center
    cross_pos = {(1, 0): 0, (-1, 0): 1, (0, 1): 2, (0, -1): 3, (2, 0): 4, (-2, 0): 5, (0, 2): 6, (0, -2): 7}
    for i in range(N):
        c = L["node"][i]
        d = model.size[c]
        for b in range(-2, 3):
            for a in range(-2, 3):
                p = (b + 2) * 5 + (a + 2)
                if a == 0 and b == 0:
                    continue
                if a == 0 or b == 0:
                    pts[i, p] = cross[i, cross_pos[(a, b)]]
                else:  # the leaf containing the off-axis point
                    q = model.locate(t["x"][c] + a * d, t["y"][c] + b * d, 99)
                    pts[i, p] = t["prim7"][q][:5]
    want = S.features(center, pts)
    got = M.bf16_bits_to_float(s.selector_features())
    assert got.shape == (N, 80) and np.all(got[:, 75:] == 0.0)
    got = got[:, :75]
    tol = 2 * 2.0 ** -8 * np.abs(want) + 1e-5
    bad = np.abs(got - want) > tol
    print("features: leaves", N, "entries off", int(bad.sum()), "max err", float(np.abs(got - want).max()))
    assert not bad.any()
