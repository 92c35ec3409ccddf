"""The benchmark configurations at their stated shapes on the GPU (through the C-ABI).

c1 (vortex, 64x64, level 2, 100 steps) runs against the oracle in test_gpu_parity.py. Here:

* c2: seeded 4-quadrant Riemann problem, base 256x256, level 4, regrid every 10 steps: the
  FP32-flux / FP64-update path against the plain FP64 path for 100 steps;
* c3: Mach-2 shock and one seeded bubble, base 512x256, level 5: FP32 flux against FP64 through
  a regrid;
* c5: Mach-3 shock through four seeded bubbles, base 1024x512, level 4, the benchmark's
  developed 18.9 M-leaf mesh: the learned selector + FP32 flux against FP64.

Both paths always run on identical meshes: at a regrid the FP64 context's sensors decide for
both (fdmamr_regrid_sensors replays them in the FP32 context), so refinement, coarsening and
balancing act the same way on each context's own state. Accuracy bar: L1(rho) <= 1e-4 (area
weighted, output_ADT.F90:43-61), positive density and pressure.
"""
import argparse

import numpy as np
import pytest

import fdmamr as F
from helpers import riemann_c2

pytestmark = pytest.mark.gpu

def same_mesh_clone(src, cfg):
    t = src.tree()
    m, n = int(t["num"][:, 1].max()), int(t["num"][:, 0].max())
    c = F.Solver(cfg)
    c.import_tree(m, n, t)
    c.build_tables()
    return c


def regrid_pair(s64, s32):
    """Regrid both contexts with the FP64 context's sensors: identical meshes afterwards."""
    rot, div = s64.sensors()
    s64.regrid()
    s32.regrid_with_sensors(rot, div)
    assert np.array_equal(s32.geometry()["node"], s64.geometry()["node"])
    return s32


def l1_rho(s, a, b):
    lvl = s.geometry()["lvl"]
    area = (1.0 / 2.0 ** lvl) ** 2
    return float(np.sum(np.abs(a[:, 0] - b[:, 0]) * area) / np.sum(area))


def test_c2_riemann_256_level4_fp32_vs_fp64_100_steps():
    cfg64 = riemann_c2()  # base 256x256, level 4, density-gradient flags, regrid every 10
    s64 = F.Solver(cfg64).setup()
    s32 = same_mesh_clone(s64, riemann_c2(flux_precision=F.FLUX_FP32))
    s32.set_state(s64.get_state())
    n0 = s64.n_leaf
    print("c2 leaves", n0, "l_max", s64.counts()["l_max"])
    assert s64.counts()["l_max"] == 4 and n0 > 100000
    for r in range(10):
        s64.run_steps(10)
        s32.run_steps(10)
        if r < 9:
            s32 = regrid_pair(s64, s32)
    a, b = s64.get_state(), s32.get_state()
    err = l1_rho(s64, a, b)
    print("c2: 100 steps, 9 regrids, t =", s64.tt, "leaves", s64.n_leaf, "L1(rho) FP32 vs FP64", err)
    assert err <= 1e-4
    assert s32.check_positivity() == 0 and s64.check_positivity() == 0


def c3_config(**kw):
    g, M = 1.4, 2.0
    p2 = 1 + 2 * g / (g + 1) * (M * M - 1)
    r2 = (g + 1) * M * M / ((g - 1) * M * M + 2)
    u2 = M * np.sqrt(g) * (1 - 1 / r2)
    d = dict(xmin=0.0, xmax=2.0, ymin=0.0, ymax=1.0, h=1.0 / 256, gamma=g,
             bc=(F.BC_INFLOW, F.BC_TRANSMISSIVE, F.BC_TRANSMISSIVE, F.BC_TRANSMISSIVE),
             inflow=(r2, u2, 0.0, p2), ic_kind=F.IC_SHOCK_BUBBLES, seed=2,
             ic_param=(0.4, M, 1, 0.2, 0.2, 0.1, 5.0), localdt=1, cfl=0.5, t_end=1e9, amr=1, lvl_amr=5,
             pstep_adapt=10, amr_sensor=1, pre_adapt=5, wls_local=1)
    d.update(kw)
    return F.Config(**d)


def test_c3_shock_bubble_512x256_level5_fp32_vs_fp64_through_regrid():
    s64 = F.Solver(c3_config()).setup()
    s32 = same_mesh_clone(s64, c3_config(flux_precision=F.FLUX_FP32))
    s32.set_state(s64.get_state())
    print("c3 leaves", s64.n_leaf, "l_max", s64.counts()["l_max"])
    assert s64.counts()["l_max"] == 5
    for r in range(2):
        s64.run_steps(10)
        s32.run_steps(10)
        if r == 0:
            s32 = regrid_pair(s64, s32)
    err = l1_rho(s64, s64.get_state(), s32.get_state())
    print("c3: 20 steps, 1 regrid, leaves", s64.n_leaf, "L1(rho) FP32 vs FP64", err)
    assert err <= 1e-4
    assert s32.check_positivity() == 0 and s64.check_positivity() == 0


def test_c5_shock4bubbles_learned_vs_fp64():
    # the benchmark's measured mesh (developed flow, 1024x512 base, level 4, ~19 M leaves, right
    # after a regrid), the selector pre-trained on the vortex and Riemann configurations, 2%
    # judge per step
    import bench

    args = argparse.Namespace(workload="shock4bubbles", refine_goal=0.03, develop_t=0.4, develop_lvl=1,
                              refine_passes=3, spinup_cycles=2, regrid_every=10)
    w, _ = bench.pretrain_selector(7, rows_target=2_000_000, iters=600, threshold=0.5, label_bound=0.05)
    s64 = bench.build_mesh(args, F)
    cnt = s64.counts()
    print("c5 leaves", cnt["n_leaf"], "slots", cnt["n_slot_wls"])
    assert cnt["n_leaf"] > 15_000_000 and cnt["l_max"] == 4
    sel = same_mesh_clone(s64, bench.measured_config(args, F.FLUX_FP32, selector=1))
    prim0 = s64.get_state()
    sel.set_state(prim0)
    sel.selector_set_weights(w)
    sel.selector_stats(reset=True)
    s64.run_steps(10)
    sel.run_steps(10)
    st = sel.selector_stats()
    err = l1_rho(s64, s64.get_state(), sel.get_state())
    print("c5: 10 learned steps, L1(rho) vs FP64", err, "selector", st)
    assert err <= 1e-4
    assert st["judged"] > 0 and st["troubled"] < 0.5 * 10 * cnt["n_leaf"]
    assert sel.check_positivity() == 0 and s64.check_positivity() == 0
