"""CPU tests of the library's host side: C-ABI exports, mesh construction, initial conditions,
and the regrid-time flattening of the quadtree (stencil classification and WLS-WENO stencil
selection), all compared bit for bit with the oracle. These use a host-only context
(device = FDMAMR_DEVICE_HOST_ONLY), so no GPU is needed and no compute runs on the CPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
import fdmamr as F
from helpers import oracle_for, riemann_c2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fdmamr.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fdmamr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(F.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(F.EXPORTS) == syms


def host(cfg: F.Config) -> F.Solver:
    cfg.device = -1
    return F.Solver(cfg)


def test_host_only_context_refuses_device_calls():
    s = host(F.Config())
    s.setup()
    with pytest.raises(F.FdmError) as e:
        s.advance_rk3(1e-4, 1.0, 1.0)
    assert "(-5)" in str(e.value)


def test_invalid_params_rejected():
    p = F.Params()
    p.h, p.gamma, p.xmax, p.ymax, p.device = -1.0, 1.4, 1.0, 1.0, -1
    p.bc_xm = p.bc_xp = p.bc_ym = p.bc_yp = 6
    ctx = C.c_void_p()
    assert F.lib().fdmamr_create(C.byref(p), C.byref(ctx)) == -1
    p.h = 0.1
    p.bc_xm = 3  # unknown boundary kind
    assert F.lib().fdmamr_create(C.byref(p), C.byref(ctx)) == -1


def _compare_trees(a, b):
    for k in ("father", "son", "lvl", "sp", "num", "cross"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"])
    assert np.array_equal(a["prim7"][:, :5], b["prim7"][:, :5])


@pytest.mark.parametrize("kind", ["vortex", "riemann", "shock"])
def test_mesh_and_initial_condition_match_oracle(kind):
    if kind == "vortex":
        cfg = F.Config(h=10.0 / 40)
    elif kind == "riemann":
        cfg = riemann_c2(h=1.0 / 40, pre_adapt=0)
    else:
        cfg = F.Config(xmin=0, xmax=2, ymin=0, ymax=1, h=1 / 32, gamma=1.4,
                       ic_kind=F.IC_SHOCK_BUBBLES, seed=5, ic_param=(0.4, 2.0, 3, 0.1, 0.2, 0.1, 5.0),
                       bc=(F.BC_INFLOW, F.BC_TRANSMISSIVE, F.BC_TRANSMISSIVE, F.BC_TRANSMISSIVE),
                       inflow=(2.0, 1.0, 0.0, 4.0))
    o = oracle_for(cfg)
    s = host(cfg).setup()
    _compare_trees(s.tree(), o.tree())


def _compare_stencils(s, o, limit=None):
    n = o.n_leaf
    kinds = np.zeros(5, int)
    idx = range(n) if limit is None else np.linspace(0, n - 1, limit).astype(int)
    for i in idx:
        si = s.stencil_info(int(i))
        for k in range(8):
            oi = o.stencil_info(int(i), k)
            kinds[oi["kind"]] += 1
            assert si["kind"][k] == oi["kind"], (i, k)
            assert si["node"][k] == oi["node"], (i, k)
            assert si["px"][k] == oi["px"] and si["py"][k] == oi["py"], (i, k)
            assert si["m1"][k] == oi["m1"] and si["m2"][k] == oi["m2"], (i, k)
            assert np.array_equal(si["nb"][k][: oi["m1"] + oi["m2"]], oi["nb"]), (i, k)
    return kinds


def test_stencil_tables_match_reference_gathers_after_amr():
    # vortex with lvlAMR = 2 after three regrids: copies, parent and child WLS-WENO slots
    o = O.Oracle(h=10.0 / 32, iflag_amr=1, lvl_amr=2, iflag_localdt=1, cfl=0.5, pstep_adapt=2)
    for _ in range(5):
        assert o.step() == 0
    s = host(F.Config(h=10.0 / 32))
    m, n = o.dims()
    s.import_tree(m, n, o.tree())
    s.build_tables()
    assert s.counts()["n_leaf"] == o.n_leaf
    g = s.geometry()
    L = o.leaves()
    assert np.array_equal(g["node"], L["node"])
    kinds = _compare_stencils(s, o)
    assert kinds[1] > 0 and kinds[3] > 0  # both WLS kinds exercised


def test_stencil_tables_deep_transmissive_mesh():
    cfg = riemann_c2(h=1.0 / 24, lvl_amr=3, pre_adapt=0)
    o = oracle_for(cfg)
    for _ in range(3):
        assert o.step() == 0
    # refine by hand through AMR steps: pstep 10 -> AMR at it=1; run once more at it=11
    s = host(cfg)
    m, n = o.dims()
    s.import_tree(m, n, o.tree())
    s.build_tables()
    _compare_stencils(s, o, limit=2000)


def test_ghost_aliases_and_fast_fill_items():
    # Ghosts that copy one interior leaf (and every inflow ghost but the first) are read through
    # that leaf; the fast path refills only the others per stage. On this transmissive mesh the
    # refinement reaches the boundary, so some ghosts are restriction means of refined partners
    # and some WLS slots read them.
    cfg = riemann_c2(h=1.0 / 24, lvl_amr=3, pre_adapt=0)
    o = oracle_for(cfg)
    for _ in range(3):
        assert o.step() == 0
    s = host(cfg)
    m, n = o.dims()
    s.import_tree(m, n, o.tree())
    s.build_tables()
    c, ti = s.counts(), s.table_info()
    print(c, ti)
    assert ti["slots"] == c["n_slot_wls"]
    assert ti["ghosts_aliased"] + ti["fill_items_fast"] - c["n_slot_restrict"] == c["n_ghost"]
    assert 0 < ti["ghosts_aliased"] < c["n_ghost"]
    assert 0 < ti["slots_plain"] < ti["slots"]
    assert ti["slots_generic"] == 0
    # the classification the tables were built from is unchanged (node ids, not state slots)
    _compare_stencils(s, o, limit=500)


def test_sixteen_byte_records_escape_only_far_entries():
    # The fast path's 16-byte stencil records hold a neighbour leaf as its offset from the leaf
    # (below 16 384) and a ghost or slot as an offset from its block's base; on a uniform
    # 256 x 256 mesh the y-neighbours across the boundaries of the largest Morton blocks of base
    # cells lie tens of thousands of leaves away, so those leaves (and only a few others) are
    # read from the 32-byte record. The table build decodes every record and checks it against
    # the 32-byte one (it raises otherwise).
    s = host(riemann_c2(h=1.0 / 256, lvl_amr=0, pre_adapt=0))
    s.setup()
    n = s.counts()["n_leaf"]
    esc = s.table_info()["records16_escaped"]
    print("leaves", n, "escaped 16-byte records", esc)
    assert n == 256 * 256
    assert 0 < esc < 0.05 * n
    # an AMR mesh with level jumps (slot ids encoded from the blocks' bases)
    cfg = riemann_c2(h=1.0 / 24, lvl_amr=3, pre_adapt=0)
    o = oracle_for(cfg)
    for _ in range(3):
        assert o.step() == 0
    s2 = host(cfg)
    m, n2 = o.dims()
    s2.import_tree(m, n2, o.tree())
    s2.build_tables()
    ti = s2.table_info()
    assert ti["slots"] > 0 and ti["records16_escaped"] < 0.05 * s2.counts()["n_leaf"]


def test_tree_round_trip_and_canonical_ids():
    o = O.Oracle(h=10.0 / 32, iflag_amr=1, lvl_amr=2, pstep_adapt=2, iflag_localdt=1, cfl=0.5)
    for _ in range(3):
        assert o.step() == 0
    t = o.tree()
    s = host(F.Config(h=10.0 / 32))
    m, n = o.dims()
    s.import_tree(m, n, t)
    _compare_trees(s.tree(), t)


def _product_tables_equal_oracle(o, cfg, limit=None):
    s = host(cfg)
    m, n = o.dims()
    s.import_tree(m, n, o.tree())
    s.build_tables()
    assert s.counts()["n_leaf"] == o.n_leaf
    assert np.array_equal(s.geometry()["node"], o.leaves()["node"])
    return _compare_stencils(s, o, limit=limit)


def test_stencil_tables_periodic_mesh_refined_at_the_boundary():
    # the periodic density-wave mesh whose level-2 ghosts take the reversed son path
    # (pinned in test_oracle_stencil_pins.py); the product's tables equal the oracle's there
    from test_oracle_stencil_pins import periodic_wave_kwargs
    kw = periodic_wave_kwargs()
    o = O.Oracle(**kw)
    for _ in range(7):
        assert o.step() == 0
    cfg = F.Config(xmin=0.0, xmax=2.0, ymin=0.0, ymax=2.0, h=2 / 24, wls_local=1)
    kinds = _product_tables_equal_oracle(o, cfg, limit=3000)
    assert kinds[4] > 0 and kinds[3] > 0 and kinds[1] > 0


@pytest.mark.parametrize("refine", [
    [(4.5, 4.5, 0), (4.25, 4.75, 1)],                                       # child m = 7
    [(4.5, 4.5, 0), (4.25, 4.75, 1), (4.5, 3.5, 0)],                        # child m = 8
    [(0.5, 4.5, 0), (-0.5, 5.5, 0), (-0.25, 5.25, 1)],                      # child m = 6 (copy)
    [(4.5, 4.5, 0), (4.25, 4.25, 1), (4.125, 4.125, 2)],                    # parent m = 6 (interpA)
    [(4.5, 4.5, 0), (4.25, 4.25, 1), (4.125, 4.125, 2), (4.5, 3.5, 0)],     # parent m = 7
])
def test_stencil_tables_hand_built_trees(refine):
    # the hand-built trees whose stencil lists were derived by hand (test_oracle_stencil_pins.py)
    from test_oracle_stencil_pins import _hand_tree
    o = _hand_tree(refine)
    kinds = _product_tables_equal_oracle(o, F.Config(h=1.0))
    assert kinds[1] + kinds[2] + kinds[3] + kinds[4] > 0
