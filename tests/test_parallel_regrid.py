"""The host regrid on several threads against the same regrid on one thread and against the
oracle's serial regrid (orc_regrid_sensors). On many threads the library runs solute_ref over
the base cells in parallel and gridmodify1 / gridmodify4 row by row over the base cells, takes
son blocks from a shared pool and renumbers the nodes canonically afterwards; the trees must
come out node for node (canonical numbering, with the leaves' values) as the serial loops make
them (mesh.F90:305-395, solution.F90:468-556). Host-only contexts with caller-supplied sensors:
no GPU is used."""
import numpy as np
import pytest

import fdmamr as F
from helpers import oracle_for
from test_host_tables import _compare_stencils, _compare_trees, host


def _cfg(lvl_amr=3):
    # shock-bubble domain on a 134 x 70 base grid (with ghosts): enough base cells for the
    # parallel passes
    return F.Config(xmin=0.0, xmax=2.0, ymin=0.0, ymax=1.0, h=1.0 / 64, gamma=1.4,
                    ic_kind=F.IC_SHOCK_BUBBLES, seed=3, ic_param=(0.4, 2.0, 3, 0.1, 0.2, 0.1, 5.0),
                    bc=(F.BC_INFLOW, F.BC_TRANSMISSIVE, F.BC_TRANSMISSIVE, F.BC_TRANSMISSIVE),
                    inflow=(2.0, 1.0, 0.0, 4.0), amr=1, lvl_amr=lvl_amr, wls_local=1,
                    refine_goal_rot=1.0, refine_goal_div=1.0, coarse_goal_rot=0.4, coarse_goal_div=0.4)


def _sensors(x, y, k):
    """Rings and a wavy line (moving with k) times a hashed per-position factor, so that the
    refinement is patchy and both refines and coarsens, with many level jumps."""
    s = np.zeros_like(x)
    for r in range(4):
        cx, cy = 0.3 + 0.45 * r + 0.05 * k, 0.5 + 0.25 * np.sin(1.3 * r + k)
        d = np.abs(np.hypot(x - cx, y - cy) - 0.15 - 0.02 * r)
        s += np.exp(-d * d / 2e-3)
    d = np.abs(x - 1.6 + 0.1 * k - 0.05 * np.sin(6 * y))
    s += np.exp(-d * d / 1e-3)
    h = (np.floor(x * 4096).astype(np.int64) * 73856093) ^ (np.floor(y * 4096).astype(np.int64) * 19349663)
    nz = ((h ^ (h >> 13)) & 0xFFFF) / 65536.0
    s = s * (0.3 + nz)
    return s, 0.7 * s + 0.1 * nz


def test_parallel_regrid_matches_serial_and_oracle(monkeypatch):
    cfg = _cfg()
    o = oracle_for(cfg)
    ser = host(cfg).setup()
    par = host(cfg).setup()
    counts = []
    for k in range(4):
        L = o.leaves()
        rot, div = _sensors(L["x"], L["y"], k)
        o.regrid_sensors(rot, div)
        monkeypatch.setenv("FDMAMR_HOST_THREADS", "1")
        ser.regrid_with_sensors(rot, div)
        monkeypatch.setenv("FDMAMR_HOST_THREADS", "6")
        par.regrid_with_sensors(rot, div)
        t_o, t_s, t_p = o.tree(), ser.tree(), par.tree()
        _compare_trees(t_s, t_o)
        _compare_trees(t_p, t_o)
        assert np.array_equal(t_p["prim7"], t_s["prim7"])
        counts.append(o.n_leaf)
        assert par.counts()["n_leaf"] == o.n_leaf
    lv = o.leaves()["lvl"]
    print("leaves per regrid", counts, "levels", np.bincount(lv))
    assert len(set(counts)) == len(counts)          # every regrid changed the mesh
    assert set(np.unique(lv)) == {0, 1, 2, 3}
    kinds = _compare_stencils(par, o, limit=4000)
    assert kinds[1] > 0 and kinds[3] > 0


def test_parallel_regrid_any_thread_count(monkeypatch):
    # the row-pipelined balance passes and the skipping of settled cells must give the
    # one-thread tree for any number of workers (odd counts, and more workers than cores)
    cfg = _cfg(lvl_amr=4)
    cfg.h = 1.0 / 96
    ctx = {nt: host(cfg).setup() for nt in (1, 2, 3, 16)}
    for k in range(5):
        g = ctx[1].geometry()
        rot, div = _sensors(g["x"], g["y"], 0.7 * k)
        trees = {}
        for nt, s in ctx.items():
            monkeypatch.setenv("FDMAMR_HOST_THREADS", str(nt))
            s.regrid_with_sensors(rot, div)
            trees[nt] = s.tree()
        for nt in (2, 3, 16):
            for key in F.Solver.TREE_KEYS:
                assert np.array_equal(trees[nt][key], trees[1][key]), (k, nt, key)
    assert ctx[1].counts()["l_max"] == 4
