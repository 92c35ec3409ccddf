"""Checkpoint and data dumps (Solver.save / restore, write_selector_rows).

The reference keeps its state in memory and has no restart file (it writes only Residual.dat
and accuracy_error.dat, output_ADT.F90:7-75); a checkpoint here is the quadtree with U(1:7)
of every node (what fdmamr_tree_export gives), the driver's scalars and, for a learned
context, the network weights, the per-leaf flags and the judge's sample counter. A run resumed
from a checkpoint must continue bit for bit as the uninterrupted run, on every path, through
a regrid. The selector rows dump holds the network's BF16 input rows and the labels.
"""
import numpy as np
import pytest

import fdmamr as F
from helpers import riemann_c2
from test_host_tables import host
from test_parallel_regrid import _cfg, _sensors


def test_checkpoint_round_trip_host_only(tmp_path):
    # an AMR tree made by host regrids with given sensors; the restored context has the same
    # tree node for node, the same leaf tables and the driver's scalars
    a = host(_cfg()).setup()
    for k in range(2):
        g = a.geometry()
        a.regrid_with_sensors(*_sensors(g["x"], g["y"], k))
    a.tt, a.it, a.tstep, a.lam = 0.125, 17, 3.5e-4, np.array([2.5, 1.25])
    p = tmp_path / "ck.npz"
    a.save(p)
    b = host(_cfg()).restore(p)
    ta, tb = a.tree(), b.tree()
    for k in F.Solver.TREE_KEYS:
        assert np.array_equal(ta[k], tb[k]), k
    assert a.counts() == b.counts()
    ga, gb = a.geometry(), b.geometry()
    for k in ("x", "y", "lvl", "node"):
        assert np.array_equal(ga[k], gb[k]), k
    assert (b.tt, b.it, b.tstep) == (0.125, 17, 3.5e-4) and np.array_equal(b.lam, [2.5, 1.25])
    # the file holds plain arrays only
    with np.load(p, allow_pickle=False) as z:
        assert set(F.Solver.TREE_KEYS) <= set(z.files)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "fp32", "learned"])
def test_resumed_run_matches_the_uninterrupted_run(mode, tmp_path):
    kw = dict(h=1.0 / 64, lvl_amr=2, pre_adapt=2)
    if mode != "fp64":
        kw["flux_precision"] = F.FLUX_FP32
    if mode == "learned":
        kw["selector"] = 1
    cfg = riemann_c2(**kw)
    a = F.Solver(cfg).setup()
    if mode == "learned":
        a.selector_init(5)
        a.selector_config(0.5, 0.25, 0.05)  # a large judge sample, so that it matters
    a.run_steps(4)
    a.regrid()
    a.run_steps(3)
    p = tmp_path / "ck.npz"
    a.save(p)
    b = F.Solver(cfg).restore(p)
    if mode == "learned":
        b.selector_config(0.5, 0.25, 0.05)
        assert np.array_equal(a.get_flags(), b.get_flags())
    for _ in range(2):
        a.run_steps(3)
        b.run_steps(3)
        assert a.tt == b.tt
        assert np.array_equal(a.get_state(), b.get_state())
        if mode == "learned":
            assert np.array_equal(a.get_flags(), b.get_flags())
        a.regrid()
        b.regrid()
        assert np.array_equal(a.geometry()["node"], b.geometry()["node"])
    print(mode, "leaves", a.n_leaf, "t", a.tt)


@pytest.mark.gpu
def test_selector_rows_dump(tmp_path):
    s = F.Solver(riemann_c2(h=1.0 / 64, lvl_amr=2, pre_adapt=2, flux_precision=F.FLUX_FP32,
                            selector=1)).setup()
    s.run_steps(3)
    p = tmp_path / "rows.npz"
    s.write_selector_rows(p)
    d = F.Solver.read_selector_rows(p)
    assert d["features"].dtype == np.uint16 and d["features"].shape == (s.n_leaf, F.SELECTOR_NFEAT)
    assert np.array_equal(d["features"], s.selector_features())
    assert np.array_equal(d["labels"], s.labels())
    g = s.geometry()
    assert np.array_equal(d["x"], g["x"]) and np.array_equal(d["lvl"], g["lvl"])
    # the dumped rows train like the device-resident set: the gradient of the dumped rows
    # equals the gradient of the same rows passed directly
    s.selector_init(2)
    y = (d["labels"] > 0.05).astype(np.uint8)
    l1, g1 = s.selector_grad(d["features"][:4096], y[:4096])
    l2, g2 = s.selector_grad(s.selector_features()[:4096], y[:4096])
    assert l1 == l2 and np.array_equal(g1, g2)
