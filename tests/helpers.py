"""Shared helpers for the test suite: oracle/product set-up on identical meshes."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
import fdmamr as F


def oracle_kwargs_from_config(cfg: F.Config) -> dict:
    return dict(xmin=cfg.xmin, xmax=cfg.xmax, ymin=cfg.ymin, ymax=cfg.ymax, h=cfg.h,
                gamma=cfg.gamma, Ma=cfg.Ma, tstep=cfg.tstep, cfl=cfg.cfl, t_end=cfg.t_end,
                iflag_localdt=cfg.localdt, iflag_amr=cfg.amr, lvl_amr=cfg.lvl_amr,
                pstep_adapt=cfg.pstep_adapt, nstep_adapt=cfg.nstep_adapt,
                refine_goal_rot=cfg.refine_goal_rot, refine_goal_div=cfg.refine_goal_div,
                coarse_goal_rot=cfg.coarse_goal_rot, coarse_goal_div=cfg.coarse_goal_div,
                amr_sensor=cfg.amr_sensor, bc_xm=cfg.bc[0], bc_xp=cfg.bc[1], bc_ym=cfg.bc[2],
                bc_yp=cfg.bc[3], inflow=cfg.inflow, ic_kind=cfg.ic_kind, seed=cfg.seed,
                ic_param=tuple(cfg.ic_param) + (0.0,) * (8 - len(cfg.ic_param)),
                pre_adapt=cfg.pre_adapt, wls_local=cfg.wls_local)


def oracle_for(cfg: F.Config) -> O.Oracle:
    return O.Oracle(**oracle_kwargs_from_config(cfg))


def product_from_oracle(o: O.Oracle, cfg: F.Config) -> F.Solver:
    """A product context on exactly the oracle's current tree and state."""
    s = F.Solver(cfg)
    m, n = o.dims()
    s.import_tree(m, n, o.tree())
    s.build_tables()
    return s


def rel_err(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / np.maximum(den, 1.0))) if a.size else 0.0


def riemann_c2(h=1.0 / 256, **kw) -> F.Config:
    d = dict(xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, h=h, gamma=1.4,
             bc=(F.BC_TRANSMISSIVE,) * 4, ic_kind=F.IC_RIEMANN4, seed=1,
             ic_param=(0.5, 0.5), localdt=1, cfl=0.5, t_end=0.2, amr=1, lvl_amr=4,
             pstep_adapt=10, amr_sensor=1, pre_adapt=4, wls_local=1)
    d.update(kw)
    return F.Config(**d)
