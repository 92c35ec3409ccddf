"""Failure paths: the reference's NaN stop (solution.F90:209-215), the singular branch of brinv
(XuanGua.F90:616-619), the FP64 positivity check of the step (new), the judge forcing
cleared leaves back to the full path (new), and the C-ABI's rejection of malformed trees.

CPU tests use the oracle and host-only contexts; GPU tests drive the library's error codes
(-3 NaN residual, -4 non-positive density or pressure) through the C-ABI.
"""
import numpy as np
import pytest

from oracle import oracle as O
import fdmamr as F
from helpers import riemann_c2


# ---------------------------------------------------------------- CPU: oracle
def test_oracle_nan_residual_stops():
    # a NaN density reaches the RMS residual; the reference stops there (solution.F90:211-213)
    o = O.Oracle()
    p = o.leaves()["prim"]
    p[37, 0] = np.nan
    o.set_leaves(p)
    assert o.step() == -1


def test_brinv_singular_returns_partial_matrix():
    # Gauss-Jordan with full pivoting on ones(3,3): step 1 pivots on (1,1) (the first maximum),
    # giving [[1, 1, 1], [-1, 0, 0], [-1, 0, 0]]; step 2 finds no pivot above the d + 1 == 1 test
    # and returns that partial matrix with flag 0 (XuanGua.F90:616-619, the reference prints and
    # returns it unchanged).
    flag, r = O.brinv(np.ones((3, 3)))
    assert flag == 0
    assert np.array_equal(r, np.array([[1.0, 1.0, 1.0], [-1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]]))
    # a zero matrix fails at the first pivot: the input comes back untouched
    flag, r = O.brinv(np.zeros((3, 3)))
    assert flag == 0 and np.array_equal(r, np.zeros((3, 3)))
    # rank 2: two pivots invert the non-zero block, the third finds 0
    flag, r = O.brinv(np.diag([4.0, 2.0, 0.0]))
    assert flag == 0 and np.array_equal(r, np.diag([0.25, 0.5, 0.0]))


# ---------------------------------------------------------------- CPU: tree import
def _host(cfg):
    cfg.device = -1
    return F.Solver(cfg)


def _tree_copy(t):
    return {k: np.array(v, copy=True) for k, v in t.items()}


def test_tree_import_rejects_malformed_trees_and_keeps_the_old_one():
    o = O.Oracle(h=10.0 / 32, iflag_amr=1, lvl_amr=2, pstep_adapt=2, iflag_localdt=1, cfl=0.5)
    for _ in range(3):
        assert o.step() == 0
    t = o.tree()
    m, n = o.dims()
    s = _host(F.Config(h=10.0 / 32))
    s.import_tree(m, n, t)
    s.build_tables()
    ref = s.tree()
    refined = int(np.flatnonzero(t["son"] >= 0)[0])
    first_son = int(t["son"][refined])
    leaf = int(np.flatnonzero((t["son"] < 0) & (t["father"] >= 0))[0])
    cases = {
        "son past the end": lambda u: u["son"].__setitem__(refined, len(u["son"]) - 2),
        "son block into base cells": lambda u: u["son"].__setitem__(refined, 5),
        "sp wrong": lambda u: u["sp"].__setitem__(first_son, 3),
        "level wrong": lambda u: u["lvl"].__setitem__(leaf, u["lvl"][leaf] + 1),
        "father out of range": lambda u: u["father"].__setitem__(leaf, -7),
        "base cell with a father": lambda u: u["father"].__setitem__(0, 1),
        "cross out of range": lambda u: u["cross"].__setitem__(3, 9),
        "base index wrong": lambda u: u["num"].__setitem__((0, 1), 2),
        "centre not finite": lambda u: u["x"].__setitem__(leaf, np.nan),
        "too deep": lambda u: u["lvl"].__setitem__(slice(m * n, None), 40),
    }
    for name, corrupt in cases.items():
        u = _tree_copy(t)
        corrupt(u)
        with pytest.raises(F.FdmError) as e:
            s.import_tree(m, n, u)
        assert "(-1)" in str(e.value), name
        # the context keeps its old tree; its tables must be rebuilt before use
        with pytest.raises(F.FdmError):
            s.stencil_info(0)
        s.build_tables()
        back = s.tree()
        for k in ("father", "son", "lvl", "sp", "num", "cross"):
            assert np.array_equal(back[k], ref[k]), (name, k)


# ---------------------------------------------------------------- GPU: error codes of the step
@pytest.mark.gpu
def test_nonpositive_pressure_in_a_stage_returns_minus4():
    # dt = 1.0 on the reference's default vortex is ~25x the CFL limit: the first stage's
    # forward-Euler substep drives the pressure negative (finite values), so the step reports
    # the positivity failure (-4) rather than the NaN it leads to afterwards.
    o = O.Oracle()
    s1 = o.substep(1.0, 2.9, 2.9)
    assert np.isfinite(s1).all() and s1[:, 4].min() < 0.0
    for prec in (F.FLUX_FP64, F.FLUX_FP32):
        s = F.Solver(F.Config(tstep=1.0, flux_precision=prec)).setup()
        with pytest.raises(F.FdmError) as e:
            s.run_steps(1)
        assert "(-4)" in str(e.value), prec
        s.close()
        s = F.Solver(F.Config(flux_precision=prec)).setup()
        with pytest.raises(F.FdmError) as e:
            s.advance_rk3(1.0, 2.9, 2.9)
        assert "(-4)" in str(e.value), prec
        s.close()
    # a CFL-limited step of the same state passes
    s = F.Solver(F.Config(localdt=1, cfl=0.5, t_end=1e9)).setup()
    s.run_steps(3)
    assert s.check_positivity() == 0


@pytest.mark.gpu
def test_nan_state_returns_minus3_and_positivity_counts():
    s = F.Solver(F.Config()).setup()
    p = s.get_state()
    p[37, 0] = np.nan
    s.set_state(p)
    with pytest.raises(F.FdmError) as e:
        s.run_steps(1)
    assert "(-3)" in str(e.value)
    s.close()
    s = F.Solver(F.Config()).setup()
    p = s.get_state()
    p[5, 4] = -0.25
    p[9, 0] = 0.0
    s.set_state(p)
    assert s.check_positivity() == 2


@pytest.mark.gpu
def test_all_clear_selector_is_forced_back_by_the_judge():
    # A network that clears every leaf (sigmoid(-30) ~ 0) with the judge checking every cleared
    # leaf (fraction 1.0): after one step exactly the leaves whose label exceeds the bound are
    # forced back to the full path (flag bit 2), and they take it in the next step.
    s = F.Solver(riemann_c2(h=1.0 / 64, lvl_amr=2, pre_adapt=2, flux_precision=F.FLUX_FP32,
                            selector=1)).setup()
    w = np.zeros(F.SELECTOR_NPARAM, np.float32)
    w[-1] = -30.0
    s.selector_set_weights(w)
    bound = 0.05
    # the leaves the set-up's regrids made stay troubled through the selector's first decision:
    # one step without a judge sample uses that up
    s.selector_config(0.5, 0.0, bound)
    s.run_steps(1)
    s.selector_stats(reset=True)
    s.selector_config(0.5, 1.0, bound)
    lab = s.labels()
    s.run_steps(1)
    st = s.selector_stats(reset=True)
    flags = s.get_flags()
    forced = (flags & 4) != 0
    want = lab > bound
    # labels within 1e-4 of the bound may round either way (the step's slot launch forms
    # ghost points from their entry lists in FP32)
    sure = np.abs(lab - bound) > 1e-4 * bound
    print("judged", st["judged"], "rejected", st["rejected"], "labelled troubled", int(want.sum()))
    assert st["judged"] == s.n_leaf
    assert want.sum() > 100
    assert np.array_equal(forced[sure], want[sure])
    assert st["rejected"] == int(forced.sum())
    # with no judge sample the next step puts exactly the forced leaves on the full path
    s.selector_config(0.5, 0.0, bound)
    s.run_steps(1)
    st2 = s.selector_stats(reset=True)
    cleared, troubled = s.leaf_lists()
    assert np.array_equal(troubled, np.flatnonzero(forced))
    assert st2["troubled"] == int(forced.sum()) and st2["judged"] == 0
