"""ctypes binding of libfdmamr (include/fdmamr.h) and a small driver mirroring the reference.

`Solver` follows the reference driver's call order (MAIN.F90:8-22 and solute,
solution.F90:171-238): creat_mesh -> initial -> IDgrid/table flattening -> the step loop
(lambda, dt, Time_advance_RK, residual) with AMR every `pstep_adapt` steps. Every numerical
step runs in the HIP library; this module only moves arguments. The library must be built
(build.py); loading fails loudly otherwise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libfdmamr.so")

BC_TRANSMISSIVE, BC_INFLOW, BC_PERIODIC = 1, 2, 6
FLUX_FP64, FLUX_FP32 = 0, 1
IC_VORTEX, IC_RIEMANN4, IC_SHOCK_BUBBLES, IC_DENSITY_WAVE = 0, 1, 2, 3
GAMMA_REF = float(np.float32(1.4))  # the reference's REAL(4) literal (MAIN.F90:38)

# every symbol include/fdmamr.h declares
EXPORTS = (
    "fdmamr_create", "fdmamr_destroy", "fdmamr_last_error", "fdmamr_creat_mesh",
    "fdmamr_initial", "fdmamr_tree_import", "fdmamr_tree_export", "fdmamr_tree_size",
    "fdmamr_build_tables", "fdmamr_counts", "fdmamr_set_state", "fdmamr_get_state",
    "fdmamr_get_ghost_state", "fdmamr_leaf_geometry", "fdmamr_max_wavespeed",
    "fdmamr_max_wavespeed_noalloc",
    "fdmamr_advance_rk3", "fdmamr_run_steps", "fdmamr_check_positivity", "fdmamr_sensors",
    "fdmamr_regrid", "fdmamr_regrid_sensors", "fdmamr_stencil_info", "fdmamr_stencil_values", "fdmamr_stencil_values_fast", "fdmamr_table_info", "fdmamr_leaf_lists", "fdmamr_set_slot_merge", "fdmamr_set_face_share", "fdmamr_profile",
    "fdmamr_profile_read", "fdmamr_profile_read_select", "fdmamr_selector_set_flags", "fdmamr_selector_get_flags",
    "fdmamr_labels", "fdmamr_selector_init", "fdmamr_selector_set_weights",
    "fdmamr_selector_get_weights", "fdmamr_selector_config", "fdmamr_selector_infer",
    "fdmamr_selector_features", "fdmamr_selector_collect", "fdmamr_selector_clear_dataset",
    "fdmamr_selector_train", "fdmamr_selector_grad", "fdmamr_selector_stats",
    "fdmamr_selector_judge_counter",
)
SELECTOR_NPARAM = 13569
SELECTOR_NFEAT = 80


class Params(C.Structure):
    _fields_ = [
        ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
        ("h", C.c_double), ("gamma", C.c_double), ("Ma", C.c_double),
        ("bc_xm", C.c_int32), ("bc_xp", C.c_int32), ("bc_ym", C.c_int32), ("bc_yp", C.c_int32),
        ("inflow", C.c_double * 4),
        ("flux_precision", C.c_int32), ("selector", C.c_int32), ("device", C.c_int32),
        ("wls_local", C.c_int32),
    ]


class FdmError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libfdmamr.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FdmError(f"{LIB_PATH} not built: run python fdm-cell-based-amr-cartesian_amd/build.py")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    sig = {
        "fdmamr_create": ([C.POINTER(Params), C.POINTER(P)], C.c_int32),
        "fdmamr_destroy": ([P], None),
        "fdmamr_last_error": ([P, C.c_char_p, C.c_int32], C.c_int32),
        "fdmamr_creat_mesh": ([P], C.c_int32),
        "fdmamr_initial": ([P, C.c_int32, C.c_uint64, dp], C.c_int32),
        "fdmamr_tree_import": ([P, C.c_int32, C.c_int32, C.c_int32, ip, ip, ip, ip, ip, ip, dp, dp, dp], C.c_int32),
        "fdmamr_tree_export": ([P, ip, ip, ip, ip, ip, ip, ip, dp, dp, dp], C.c_int32),
        "fdmamr_tree_size": ([P, ip], C.c_int32),
        "fdmamr_build_tables": ([P], C.c_int32),
        "fdmamr_counts": ([P, ip, ip, ip, ip, ip], C.c_int32),
        "fdmamr_set_state": ([P, dp], C.c_int32),
        "fdmamr_get_state": ([P, dp], C.c_int32),
        "fdmamr_get_ghost_state": ([P, dp], C.c_int32),
        "fdmamr_leaf_geometry": ([P, dp, dp, ip, ip], C.c_int32),
        "fdmamr_max_wavespeed": ([P, dp, dp, ip], C.c_int32),
        "fdmamr_max_wavespeed_noalloc": ([P, dp, dp, ip], C.c_int32),
        "fdmamr_advance_rk3": ([P, C.c_double, C.c_double, C.c_double, dp, dp], C.c_int32),
        "fdmamr_run_steps": ([P, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, dp, dp, dp], C.c_int32),
        "fdmamr_check_positivity": ([P, ip], C.c_int32),
        "fdmamr_sensors": ([P, dp, dp], C.c_int32),
        "fdmamr_regrid": ([P, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32], C.c_int32),
        "fdmamr_regrid_sensors": ([P, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32,
                                   dp, dp], C.c_int32),
        "fdmamr_stencil_info": ([P, C.c_int32, ip, ip, dp, dp, ip, ip, ip], C.c_int32),
        "fdmamr_stencil_values": ([P, C.c_int32, dp], C.c_int32),
        "fdmamr_table_info": ([P, ip], C.c_int32),
        "fdmamr_leaf_lists": ([P, ip, ip], C.c_int32),
        "fdmamr_set_slot_merge": ([P, C.c_int32], C.c_int32),
        "fdmamr_set_face_share": ([P, C.c_int32], C.c_int32),
        "fdmamr_stencil_values_fast": ([P, C.c_int32, C.POINTER(C.c_float)], C.c_int32),
        "fdmamr_profile": ([P, C.c_int32], C.c_int32),
        "fdmamr_profile_read": ([P, dp, ip], C.c_int32),
        "fdmamr_profile_read_select": ([P, dp, ip, C.POINTER(C.c_int64)], C.c_int32),
        "fdmamr_selector_set_flags": ([P, C.POINTER(C.c_uint8), C.c_int32], C.c_int32),
        "fdmamr_selector_get_flags": ([P, C.POINTER(C.c_uint8)], C.c_int32),
        "fdmamr_labels": ([P, C.POINTER(C.c_float)], C.c_int32),
        "fdmamr_selector_init": ([P, C.c_uint64], C.c_int32),
        "fdmamr_selector_set_weights": ([P, C.POINTER(C.c_float), C.c_int32], C.c_int32),
        "fdmamr_selector_get_weights": ([P, C.POINTER(C.c_float), C.c_int32], C.c_int32),
        "fdmamr_selector_config": ([P, C.c_float, C.c_float, C.c_float], C.c_int32),
        "fdmamr_selector_infer": ([P, C.POINTER(C.c_float), ip], C.c_int32),
        "fdmamr_selector_features": ([P, C.POINTER(C.c_uint16)], C.c_int32),
        "fdmamr_selector_collect": ([P, ip, ip], C.c_int32),
        "fdmamr_selector_clear_dataset": ([P], C.c_int32),
        "fdmamr_selector_train": ([P, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_uint64,
                                   C.POINTER(C.c_float)], C.c_int32),
        "fdmamr_selector_grad": ([P, C.POINTER(C.c_uint16), C.POINTER(C.c_uint8), C.c_int32,
                                  C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)], C.c_int32),
        "fdmamr_selector_stats": ([P, ip, ip, ip, C.c_int32], C.c_int32),
        "fdmamr_selector_judge_counter": ([P, C.c_int64, C.POINTER(C.c_uint32)], C.c_int32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


@dataclass
class Config:
    """Solver configuration; defaults follow set_parameter (MAIN.F90:29-71)."""
    xmin: float = 0.0
    xmax: float = 10.0
    ymin: float = 0.0
    ymax: float = 10.0
    h: float = 0.5
    gamma: float = GAMMA_REF
    Ma: float = 1.0
    bc: tuple = (BC_PERIODIC,) * 4          # xm, xp, ym, yp
    inflow: tuple = (0.0, 0.0, 0.0, 0.0)
    flux_precision: int = FLUX_FP64
    selector: int = 0
    device: int = 0
    ic_kind: int = IC_VORTEX
    seed: int = 0
    ic_param: tuple = (0.0,) * 8
    # time loop (solute)
    tstep: float = float(np.float32(1e-4))
    cfl: float = float(np.float32(0.45))
    t_end: float = 10.0
    localdt: int = 0
    # AMR
    amr: int = 0
    lvl_amr: int = 0
    pstep_adapt: int = 30
    nstep_adapt: int = 0
    refine_goal_rot: float = GAMMA_REF
    refine_goal_div: float = GAMMA_REF
    coarse_goal_rot: float = float(np.float32(0.4))
    coarse_goal_div: float = float(np.float32(0.4))
    amr_sensor: int = 0
    pre_adapt: int = 0
    # WLS-WENO frame: 0 absolute (the reference), 1 cell-local scaled (fine meshes)
    wls_local: int = 0


class Solver:
    """One solver context on one GPU (one HIP stream)."""

    def __init__(self, cfg: Config):
        self.L = lib()
        self.cfg = cfg
        p = Params()
        p.xmin, p.xmax, p.ymin, p.ymax = cfg.xmin, cfg.xmax, cfg.ymin, cfg.ymax
        p.h, p.gamma, p.Ma = cfg.h, cfg.gamma, cfg.Ma
        p.bc_xm, p.bc_xp, p.bc_ym, p.bc_yp = cfg.bc
        for i, v in enumerate(cfg.inflow):
            p.inflow[i] = v
        p.flux_precision, p.selector, p.device = cfg.flux_precision, cfg.selector, cfg.device
        p.wls_local = cfg.wls_local
        self.ctx = C.c_void_p()
        rc = self.L.fdmamr_create(C.byref(p), C.byref(self.ctx))
        if rc != 0:
            raise FdmError(f"fdmamr_create failed ({rc})")
        self.tt = 0.0
        self.it = 0
        self.lam = np.zeros(2)
        self.tstep = cfg.tstep
        self.res_max = np.zeros(4)
        self.res_rms = np.zeros(4)

    # -- plumbing --
    def _check(self, rc, what):
        if rc != 0:
            buf = C.create_string_buffer(512)
            self.L.fdmamr_last_error(self.ctx, buf, 512)
            raise FdmError(f"{what} failed ({rc}): {buf.value.decode()}")

    def close(self):
        if getattr(self, "ctx", None):
            self.L.fdmamr_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- driver (MAIN.F90:18-22) --
    def setup(self):
        """creat_mesh, initial, optional IC pre-adaptation, IDgrid + tables."""
        cfg = self.cfg
        self._check(self.L.fdmamr_creat_mesh(self.ctx), "creat_mesh")
        self.initial()
        self.build_tables()
        for _ in range(cfg.pre_adapt):
            self.regrid()
            self.initial()
            self.build_tables()
        return self

    def initial(self):
        cfg = self.cfg
        prm = np.array(list(cfg.ic_param) + [0.0] * (8 - len(cfg.ic_param)), np.float64)
        self._check(self.L.fdmamr_initial(self.ctx, cfg.ic_kind, cfg.seed, _dp(prm)), "initial")

    def build_tables(self):
        self._check(self.L.fdmamr_build_tables(self.ctx), "build_tables")

    def counts(self):
        a = [C.c_int32() for _ in range(5)]
        self._check(self.L.fdmamr_counts(self.ctx, *[C.byref(x) for x in a]), "counts")
        return dict(n_leaf=a[0].value, n_ghost=a[1].value, n_slot_wls=a[2].value,
                    n_slot_restrict=a[3].value, l_max=a[4].value)

    @property
    def n_leaf(self):
        return self.counts()["n_leaf"]

    # -- state --
    def get_state(self):
        out = np.zeros((self.n_leaf, 5))
        self._check(self.L.fdmamr_get_state(self.ctx, _dp(out)), "get_state")
        return out

    def get_ghost_state(self):
        out = np.zeros((self.counts()["n_ghost"], 5))
        self._check(self.L.fdmamr_get_ghost_state(self.ctx, _dp(out)), "get_ghost_state")
        return out

    def set_state(self, prim):
        prim = np.ascontiguousarray(prim, np.float64)
        self._check(self.L.fdmamr_set_state(self.ctx, _dp(prim)), "set_state")

    def geometry(self):
        n = self.n_leaf
        x = np.zeros(n); y = np.zeros(n)
        lvl = np.zeros(n, np.int32); node = np.zeros(n, np.int32)
        self._check(self.L.fdmamr_leaf_geometry(self.ctx, _dp(x), _dp(y), _ip(lvl), _ip(node)),
                    "leaf_geometry")
        return dict(x=x, y=y, lvl=lvl, node=node)

    def tree(self):
        nn = C.c_int32()
        self._check(self.L.fdmamr_tree_size(self.ctx, C.byref(nn)), "tree_size")
        n = nn.value
        t = dict(father=np.zeros(n, np.int32), son=np.zeros(n, np.int32),
                 lvl=np.zeros(n, np.int32), sp=np.zeros(n, np.int32),
                 num=np.zeros((n, 2), np.int32), cross=np.zeros(n, np.int32),
                 x=np.zeros(n), y=np.zeros(n), prim7=np.zeros((n, 7)))
        self._check(self.L.fdmamr_tree_export(self.ctx, C.byref(nn), _ip(t["father"]), _ip(t["son"]),
                                              _ip(t["lvl"]), _ip(t["sp"]), _ip(t["num"]),
                                              _ip(t["cross"]), _dp(t["x"]), _dp(t["y"]),
                                              _dp(t["prim7"])), "tree_export")
        return t

    def import_tree(self, m, n, t):
        self._check(self.L.fdmamr_tree_import(
            self.ctx, m, n, len(t["father"]), _ip(np.ascontiguousarray(t["father"], np.int32)),
            _ip(np.ascontiguousarray(t["son"], np.int32)), _ip(np.ascontiguousarray(t["lvl"], np.int32)),
            _ip(np.ascontiguousarray(t["sp"], np.int32)), _ip(np.ascontiguousarray(t["num"], np.int32)),
            _ip(np.ascontiguousarray(t["cross"], np.int32)), _dp(np.ascontiguousarray(t["x"], np.float64)),
            _dp(np.ascontiguousarray(t["y"], np.float64)),
            _dp(np.ascontiguousarray(t["prim7"], np.float64))), "tree_import")

    # -- hot path --
    def max_wavespeed(self):
        a, b, l = C.c_double(), C.c_double(), C.c_int32()
        self._check(self.L.fdmamr_max_wavespeed(self.ctx, C.byref(a), C.byref(b), C.byref(l)),
                    "max_wavespeed")
        return a.value, b.value, l.value

    def max_wavespeed_noalloc(self):
        """max_wavespeed through the library's static scratch (allocates no device memory)."""
        a, b, l = C.c_double(), C.c_double(), C.c_int32()
        self._check(self.L.fdmamr_max_wavespeed_noalloc(self.ctx, C.byref(a), C.byref(b), C.byref(l)),
                    "max_wavespeed_noalloc")
        return a.value, b.value, l.value

    def advance_rk3(self, dt, lam_x, lam_y):
        rm = np.zeros(4); rs = np.zeros(4)
        self._check(self.L.fdmamr_advance_rk3(self.ctx, dt, lam_x, lam_y, _dp(rm), _dp(rs)),
                    "advance_rk3")
        return rm, rs

    def step(self):
        """One pass of the solute loop body (solution.F90:183-230), host-synchronous."""
        cfg = self.cfg
        self.it += 1
        lx, ly, lmax = self.max_wavespeed()
        self.lam[0] = max(self.lam[0], lx)
        self.lam[1] = max(self.lam[1], ly)
        if cfg.localdt:
            p2 = float(2 ** lmax)
            self.tstep = cfg.cfl / (self.lam[0] / (1.0 * cfg.h / p2) + self.lam[1] / (cfg.h / p2))
        if self.tt + self.tstep > cfg.t_end:
            self.tstep = cfg.t_end - self.tt
        rm, rs = self.advance_rk3(self.tstep, self.lam[0], self.lam[1])
        self.tt = self.tt + self.tstep
        self.res_max = rm
        self.res_rms = np.sqrt(rs / (1.0 * self.n_leaf))
        if np.isnan(self.res_rms).any():
            raise FdmError(f"NaN residual at it={self.it}")
        if cfg.amr and self.it >= cfg.nstep_adapt and self.it % cfg.pstep_adapt == 1:
            self.regrid()

    def run_steps(self, n):
        """n steps without AMR and without host synchronisation inside (device dt)."""
        cfg = self.cfg
        tt = np.array([self.tt]); lam = self.lam.copy(); sc = np.zeros(8)
        self._check(self.L.fdmamr_run_steps(self.ctx, n, cfg.localdt, cfg.cfl, cfg.tstep, cfg.t_end,
                                            _dp(tt), _dp(lam), _dp(sc)), "run_steps")
        self.tt = float(tt[0]); self.lam = lam; self.it += n
        self.tstep = sc[1]
        self.res_rms = sc[4:8].copy()
        return sc

    def regrid(self):
        cfg = self.cfg
        self._check(self.L.fdmamr_regrid(self.ctx, cfg.lvl_amr, cfg.refine_goal_rot,
                                         cfg.refine_goal_div, cfg.coarse_goal_rot,
                                         cfg.coarse_goal_div, cfg.amr_sensor), "regrid")

    def regrid_with_sensors(self, rot, div):
        """regrid with the given per-leaf sensors (e.g. another context's on the same mesh)."""
        cfg = self.cfg
        rot = np.ascontiguousarray(rot, np.float64)
        div = np.ascontiguousarray(div, np.float64)
        assert len(rot) == len(div) == self.n_leaf
        self._check(self.L.fdmamr_regrid_sensors(self.ctx, cfg.lvl_amr, cfg.refine_goal_rot, cfg.refine_goal_div,
                                                 cfg.coarse_goal_rot, cfg.coarse_goal_div, cfg.amr_sensor,
                                                 _dp(rot), _dp(div)), "regrid_sensors")

    def sensors(self):
        n = self.n_leaf
        rot = np.zeros(n); div = np.zeros(n)
        self._check(self.L.fdmamr_sensors(self.ctx, _dp(rot), _dp(div)), "sensors")
        return rot, div

    def check_positivity(self):
        nb = C.c_int32()
        rc = self.L.fdmamr_check_positivity(self.ctx, C.byref(nb))
        if rc not in (0, -4):
            self._check(rc, "check_positivity")
        return nb.value

    # -- selector --
    def set_flags(self, flags, use=True):
        f = None if flags is None else np.ascontiguousarray(flags, np.uint8)
        ptr = None if f is None else f.ctypes.data_as(C.POINTER(C.c_uint8))
        self._check(self.L.fdmamr_selector_set_flags(self.ctx, ptr, 1 if use else 0), "set_flags")

    def get_flags(self):
        f = np.zeros(self.n_leaf, np.uint8)
        self._check(self.L.fdmamr_selector_get_flags(self.ctx, f.ctypes.data_as(C.POINTER(C.c_uint8))),
                    "get_flags")
        return f

    def labels(self):
        d = np.zeros(self.n_leaf, np.float32)
        self._check(self.L.fdmamr_labels(self.ctx, d.ctypes.data_as(C.POINTER(C.c_float))), "labels")
        return d

    def selector_init(self, seed=0):
        self._check(self.L.fdmamr_selector_init(self.ctx, seed), "selector_init")

    def selector_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(self.L.fdmamr_selector_set_weights(self.ctx, w.ctypes.data_as(C.POINTER(C.c_float)),
                                                       len(w)), "selector_set_weights")

    def selector_get_weights(self):
        w = np.zeros(SELECTOR_NPARAM, np.float32)
        self._check(self.L.fdmamr_selector_get_weights(self.ctx, w.ctypes.data_as(C.POINTER(C.c_float)),
                                                       len(w)), "selector_get_weights")
        return w

    def selector_config(self, threshold=0.5, judge_fraction=0.02, label_bound=0.05):
        self._check(self.L.fdmamr_selector_config(self.ctx, threshold, judge_fraction, label_bound),
                    "selector_config")

    def selector_infer(self):
        prob = np.zeros(self.n_leaf, np.float32)
        nt = C.c_int32()
        self._check(self.L.fdmamr_selector_infer(self.ctx, prob.ctypes.data_as(C.POINTER(C.c_float)),
                                                 C.byref(nt)), "selector_infer")
        return prob, nt.value

    def selector_features(self):
        x = np.zeros((self.n_leaf, SELECTOR_NFEAT), np.uint16)
        self._check(self.L.fdmamr_selector_features(self.ctx, x.ctypes.data_as(C.POINTER(C.c_uint16))),
                    "selector_features")
        return x

    def selector_collect(self):
        a, b = C.c_int32(), C.c_int32()
        self._check(self.L.fdmamr_selector_collect(self.ctx, C.byref(a), C.byref(b)), "selector_collect")
        return a.value, b.value

    def selector_clear_dataset(self):
        self._check(self.L.fdmamr_selector_clear_dataset(self.ctx), "selector_clear_dataset")

    def selector_train(self, iters, batch=4096, lr=2e-3, pos_weight=4.0, seed=0):
        loss = C.c_float()
        self._check(self.L.fdmamr_selector_train(self.ctx, iters, batch, lr, pos_weight, seed,
                                                 C.byref(loss)), "selector_train")
        return loss.value

    def selector_grad(self, x_bf16, y, pos_weight=1.0):
        x = np.ascontiguousarray(x_bf16, np.uint16)
        yy = np.ascontiguousarray(y, np.uint8)
        g = np.zeros(SELECTOR_NPARAM, np.float32)
        loss = C.c_float()
        self._check(self.L.fdmamr_selector_grad(self.ctx, x.ctypes.data_as(C.POINTER(C.c_uint16)),
                                                yy.ctypes.data_as(C.POINTER(C.c_uint8)), len(yy),
                                                pos_weight, g.ctypes.data_as(C.POINTER(C.c_float)),
                                                C.byref(loss)), "selector_grad")
        return loss.value, g

    def selector_stats(self, reset=False):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.L.fdmamr_selector_stats(self.ctx, C.byref(a), C.byref(b), C.byref(c),
                                                 1 if reset else 0), "selector_stats")
        return dict(troubled=a.value, judged=b.value, rejected=c.value)

    # -- measurement --
    def profile(self, enable=True):
        self._check(self.L.fdmamr_profile(self.ctx, 1 if enable else 0), "profile")

    def profile_read(self):
        ms = C.c_double(); n = C.c_int32()
        self._check(self.L.fdmamr_profile_read(self.ctx, C.byref(ms), C.byref(n)), "profile_read")
        return ms.value, n.value

    def profile_read_select(self):
        """(total ms, launches) of the selector's inference launches since the last read, and
        the 16-leaf tiles the network ran on since the selector counters were last reset."""
        ms = C.c_double(); n = C.c_int32(); tiles = C.c_int64()
        self._check(self.L.fdmamr_profile_read_select(self.ctx, C.byref(ms), C.byref(n), C.byref(tiles)),
                    "profile_read_select")
        return ms.value, n.value, tiles.value

    # -- diagnostics --
    def stencil_info(self, i):
        kind = np.zeros(8, np.int32); node = np.zeros(8, np.int32)
        px = np.zeros(8); py = np.zeros(8); m1 = np.zeros(8, np.int32); m2 = np.zeros(8, np.int32)
        nb = np.zeros(72, np.int32)
        self._check(self.L.fdmamr_stencil_info(self.ctx, i, _ip(kind), _ip(node), _dp(px), _dp(py),
                                               _ip(m1), _ip(m2), _ip(nb)), "stencil_info")
        return dict(kind=kind, node=node, px=px, py=py, m1=m1, m2=m2, nb=nb.reshape(8, 9))

    def stencil_values(self, i):
        out = np.zeros((8, 4))
        self._check(self.L.fdmamr_stencil_values(self.ctx, i, _dp(out)), "stencil_values")
        return out

    def stencil_values_fast(self, i):
        out = np.zeros((8, 4), np.float32)
        self._check(self.L.fdmamr_stencil_values_fast(self.ctx, i, out.ctypes.data_as(C.POINTER(C.c_float))),
                    "stencil_values_fast")
        return out

    def set_slot_merge(self, enable):
        """False: WLS-slot launch and ghost refill as separate launches on every mesh."""
        self._check(self.L.fdmamr_set_slot_merge(self.ctx, 1 if enable else 0), "set_slot_merge")

    def set_face_share(self, enable):
        """False: the FP64 update kernel computes every face itself (the reference's way)."""
        self._check(self.L.fdmamr_set_face_share(self.ctx, 1 if enable else 0), "set_face_share")

    def leaf_lists(self):
        """The fast path's (cleared, troubled) leaf lists as the last flagged step built them."""
        lst = np.zeros(self.n_leaf, np.int32)
        nc = np.zeros(1, np.int32)
        self._check(self.L.fdmamr_leaf_lists(self.ctx, _ip(lst), _ip(nc)), "leaf_lists")
        return lst[: int(nc[0])], lst[int(nc[0]):]

    # -- checkpoint and data dumps (no reference counterpart: the reference keeps its state in
    # memory and writes only Residual.dat and accuracy_error.dat, output_ADT.F90:7-75) --
    TREE_KEYS = ("father", "son", "lvl", "sp", "num", "cross", "x", "y", "prim7")

    def save(self, path):
        """Write a checkpoint (numpy .npz, nothing pickled): the quadtree with U(1:7) of every
        node, which holds the leaves' states; the driver's tt, it, tstep and running lambda;
        and, when the context has a selector with weights, the network weights, the per-leaf
        flags (with their sticky judge and regrid bits) and the judge's sample counter.
        restore() on a context made with the same Config continues exactly where this one
        stands (the selector's threshold, judge fraction and label bound are the caller's to
        set again with selector_config)."""
        t = self.tree()
        d = {k: t[k] for k in self.TREE_KEYS}
        d["dims"] = np.array(self.dims(t), np.int64)
        d["scalars"] = np.array([self.tt, self.tstep, self.lam[0], self.lam[1]], np.float64)
        d["it"] = np.array([self.it], np.int64)
        if self.cfg.selector:
            try:
                d["sel_w"] = self.selector_get_weights()
            except FdmError:
                pass  # no weights yet: nothing of the selector to keep
            else:
                d["sel_flags"] = self.get_flags()
                cnt = C.c_uint32()
                self._check(self.L.fdmamr_selector_judge_counter(self.ctx, -1, C.byref(cnt)), "judge_counter")
                d["sel_counter"] = np.array([cnt.value], np.int64)
        np.savez(path, **d)

    def restore(self, path):
        """Load a checkpoint written by save() into this context (made with the same Config)."""
        with np.load(path) as z:
            m, n = (int(v) for v in z["dims"])
            self.import_tree(m, n, {k: z[k] for k in self.TREE_KEYS})
            self.build_tables()
            self.tt, self.tstep, lx, ly = (float(v) for v in z["scalars"])
            self.lam = np.array([lx, ly])
            self.it = int(z["it"][0])
            if "sel_w" in z.files:
                self.selector_set_weights(z["sel_w"])
                self.set_flags(z["sel_flags"], use=True)
                self._check(self.L.fdmamr_selector_judge_counter(self.ctx, int(z["sel_counter"][0]), None),
                            "judge_counter")
        return self

    @staticmethod
    def dims(t):
        """Base grid (m columns, n rows) of an exported tree: the largest base indices."""
        return int(t["num"][:, 1].max()), int(t["num"][:, 0].max())

    def write_selector_rows(self, path):
        """Dump the selector's training rows of the current state (numpy .npz): the BF16
        feature rows (uint16 bit patterns, SELECTOR_NFEAT per leaf, the network's input as the
        device forms it), the labels (max |w1 - w1_lin| of the leaf's WCNS reconstructions)
        and the leaves' centres and levels, all in cell_NS order."""
        g = self.geometry()
        np.savez(path, features=self.selector_features(), labels=self.labels(), x=g["x"], y=g["y"],
                 lvl=g["lvl"])

    @staticmethod
    def read_selector_rows(path):
        """The arrays write_selector_rows() wrote, as a dict."""
        with np.load(path) as z:
            return {k: z[k] for k in z.files}

    def table_info(self):
        v = np.zeros(11, np.int32)
        self._check(self.L.fdmamr_table_info(self.ctx, _ip(v)), "table_info")
        return dict(slots_plain=int(v[0]), slots=int(v[1]), ghosts_aliased=int(v[2]),
                    fill_items_fast=int(v[3]), slots_generic=int(v[4]), select_tiles_direct=int(v[5]),
                    select_tiles_over64=int(v[6]), select_tile_max=int(v[7]), faces_shared_x=int(v[8]),
                    faces_shared_y=int(v[9]), records16_escaped=int(v[10]))


def config_c1(**kw) -> Config:
    """Isentropic vortex, periodic [0,10]^2, base 64x64, lvlAMR=2, CFL 0.5 local dt."""
    d = dict(h=10.0 / 64, localdt=1, cfl=0.5, amr=1, lvl_amr=2, pstep_adapt=30, t_end=1.0e9)
    d.update(kw)
    return Config(**d)
