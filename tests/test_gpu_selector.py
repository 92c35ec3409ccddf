"""GPU tests of the learned troubled-cell selector: MFMA forward and backward against the CPU
oracle, training on labels the solver produces, and an ablation showing what the trained
network gains over running without it.

Tolerances: the GPU network runs on BF16 matrix cores with FP32 accumulation; the oracle
reproduces the BF16 rounding points in FP64, so the remaining differences are FP32
accumulation order and tanh/exp in FP32: probabilities within 2e-3, gradients within 2% in
norm (and each parameter block within 5%).
"""
import numpy as np
import pytest

import fdmamr as F
from helpers import riemann_c2
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def fast_solver(**kw):
    cfg = riemann_c2(h=1.0 / 64, lvl_amr=2, pre_adapt=2, flux_precision=F.FLUX_FP32, **kw)
    return F.Solver(cfg).setup()


def test_forward_matches_oracle():
    s = fast_solver()
    w = M.random_weights(11).astype(np.float32)
    s.selector_set_weights(w)
    X = M.bf16_bits_to_float(s.selector_features())
    assert np.all(np.abs(X[:, :75]) <= 1.0 + 1e-6)
    assert np.all(X[:, 75:] == 0.0)
    prob, _ = s.selector_infer()
    want, _ = M.forward(w.astype(np.float64), X, bf16=True)
    # near-uniform stencils (features within 1e-6 of 0) are cleared without the network
    amax = np.max(np.abs(X[:, :75]), axis=1)
    gated, net = amax < 5e-7, amax > 2e-6
    assert np.all(prob[gated] == 0.0)
    assert net.sum() > 100
    err = np.max(np.abs(prob[net] - want[net]))
    print("max |p_gpu - p_oracle|", err, "gated", gated.sum(), "network", net.sum())
    assert err < 2e-3


def test_gradient_matches_oracle():
    s = fast_solver()
    w = M.random_weights(12).astype(np.float32)
    s.selector_set_weights(w)
    xb = s.selector_features()[:1000]
    rng = np.random.default_rng(0)
    y = (rng.uniform(size=len(xb)) < 0.3).astype(np.uint8)
    loss, g = s.selector_grad(xb, y, pos_weight=3.0)
    l_ref, g_ref = M.loss_and_grad(w.astype(np.float64), M.bf16_bits_to_float(xb), y, 3.0, bf16=True)
    assert abs(loss - l_ref) < 1e-3 * abs(l_ref) + 1e-5
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print("gradient relative error", rel)
    assert rel < 2e-2
    for a, b in [(M.P_W1, M.P_B1), (M.P_W2, M.P_B2), (M.P_W3, M.P_B3), (M.P_W4, M.P_B4)]:
        assert np.linalg.norm(g[a:b] - g_ref[a:b]) < 5e-2 * np.linalg.norm(g_ref[a:b]) + 1e-9


def _trained(steps_between=3, rounds=4, iters=300):
    s = fast_solver()
    s.selector_init(5)
    for _ in range(rounds):
        s.selector_collect()
        s.run_steps(steps_between)
    loss = s.selector_train(iters, batch=2048, lr=3e-3, pos_weight=4.0, seed=1)
    return s, loss


def test_training_learns_labels():
    s, loss = _trained()
    s.run_steps(3)
    lab = s.labels() > 0.05
    prob, _ = s.selector_infer()
    pred = prob >= 0.5
    recall = (pred & lab).sum() / max(lab.sum(), 1)
    cleared = 1.0 - pred.mean()
    print("loss", loss, "recall", recall, "cleared fraction", cleared, "troubled", lab.mean())
    assert recall > 0.9
    # it clears most of the leaves that really are clear
    assert cleared > 0.8 * (1.0 - lab.mean())


def test_selector_ablation_gain():
    """The trained selector keeps the full method's accuracy while clearing most leaves:
    without it, clearing every leaf gives a larger error, and taking every leaf through the
    full method costs the full method on every leaf, where the learned run sends fewer than
    60 % of the leaf-steps through it (the benchmark times the two: `value` against
    `fp32_all_troubled_value`)."""
    s, _ = _trained()
    w = s.selector_get_weights()
    cfg64 = riemann_c2(h=1.0 / 64, lvl_amr=2, pre_adapt=2)
    ref = F.Solver(cfg64).setup()
    t = ref.tree()
    m, n = int(t["num"][:, 1].max()), int(t["num"][:, 0].max())
    prim0 = ref.get_state()

    def clone(selector, flags=None):
        c = F.Solver(riemann_c2(h=1.0 / 64, lvl_amr=2, pre_adapt=2, flux_precision=F.FLUX_FP32,
                                selector=selector))
        c.import_tree(m, n, t)
        c.build_tables()
        c.set_state(prim0)
        if selector:
            c.selector_set_weights(w)
        if flags is not None:
            c.set_flags(flags, use=True)
        return c

    learned = clone(1)
    allclear = clone(0, np.zeros(len(prim0), np.uint8))
    steps = 12
    ref.run_steps(steps)
    learned.run_steps(steps)
    allclear.run_steps(steps)
    lvl = ref.geometry()["lvl"]
    area = (0.5 ** lvl) ** 2

    def l1(a):
        return float(np.sum(np.abs(a.get_state()[:, 0] - ref.get_state()[:, 0]) * area) / area.sum())

    e_learned, e_clear = l1(learned), l1(allclear)
    st = learned.selector_stats()
    print("L1 learned", e_learned, "L1 all-cleared", e_clear, "stats", st)
    assert e_learned < 1e-4
    assert e_clear > 2 * e_learned
    assert st["judged"] > 0
    assert st["troubled"] < 0.6 * steps * len(prim0)
