"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker. The product library never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BC_TRANSMISSIVE, BC_INFLOW, BC_PERIODIC = 1, 2, 6
IC_VORTEX, IC_RIEMANN4, IC_SHOCK_BUBBLES, IC_DENSITY_WAVE = 0, 1, 2, 3
GAMMA_REF = float(np.float32(1.4))  # MAIN.F90:38 (REAL(4) literal widened)


class OrcParams(C.Structure):
    _fields_ = [
        ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
        ("h", C.c_double), ("gamma", C.c_double), ("Ma", C.c_double),
        ("tstep", C.c_double), ("cfl", C.c_double), ("t_end", C.c_double),
        ("iflag_localdt", C.c_int32),
        ("iflag_amr", C.c_int32), ("lvl_amr", C.c_int32), ("pstep_adapt", C.c_int32),
        ("nstep_adapt", C.c_int32),
        ("refine_goal_rot", C.c_double), ("refine_goal_div", C.c_double),
        ("coarse_goal_rot", C.c_double), ("coarse_goal_div", C.c_double),
        ("amr_sensor", C.c_int32),
        ("bc_xm", C.c_int32), ("bc_xp", C.c_int32), ("bc_ym", C.c_int32), ("bc_yp", C.c_int32),
        ("inflow", C.c_double * 4),
        ("ic_kind", C.c_int32),
        ("seed", C.c_uint64),
        ("ic_param", C.c_double * 8),
        ("pre_adapt", C.c_int32),
        ("wls_local", C.c_int32),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.orc_new.restype = P
        L.orc_new.argtypes = [C.POINTER(OrcParams)]
        L.orc_free.argtypes = [P]
        for f in ("orc_init", "orc_step", "orc_n_leaf", "orc_n_bc", "orc_n_node"):
            getattr(L, f).argtypes = [P]
            getattr(L, f).restype = C.c_int32
        L.orc_advance_rk3.argtypes = [P, C.c_double, C.c_double, C.c_double]
        L.orc_base_dims.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.orc_get_leaves.argtypes = [P, dp, dp, ip, ip, dp]
        L.orc_get_bc_leaves.argtypes = [P, ip, dp]
        L.orc_set_leaves.argtypes = [P, dp]
        L.orc_get_scalars.argtypes = [P, dp]
        L.orc_get_residual.argtypes = [P, dp, dp]
        L.orc_tree_export.argtypes = [P, ip, ip, ip, ip, ip, ip, dp, dp, dp]
        L.orc_tree_export.restype = C.c_int32
        L.orc_stencil_info.argtypes = [P, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), ip]
        L.orc_stencil_value.argtypes = [P, C.c_int32, C.c_int32, dp]
        L.orc_sensors.argtypes = [P, dp, dp, dp]
        L.orc_sensors.restype = C.c_int32
        L.orc_wcns_z3.argtypes = [dp, dp, dp, dp, dp]
        L.orc_eigen.argtypes = [C.c_int32, dp, C.c_double, C.c_double, dp, dp]
        L.orc_flux_derivative.argtypes = [C.c_int32, dp, C.c_double, C.c_double, C.c_double,
                                          C.c_double, dp]
        L.orc_brinv.argtypes = [C.c_int32, dp, dp]
        L.orc_brinv.restype = C.c_int32
        L.orc_get_direction.argtypes = [C.c_double] * 4
        L.orc_get_direction.restype = C.c_int32
        L.orc_wls_weno.argtypes = [C.c_int32, C.c_int32, dp, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_int32, C.c_double, C.c_double, dp]
        L.orc_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_smoothness.argtypes = [C.c_int32, dp, C.c_double, C.c_double, C.c_int32,
                                     C.c_double, C.c_double]
        L.orc_smoothness.restype = C.c_double
        L.orc_neighbors.argtypes = [P, C.c_int32, ip]
        L.orc_neighbors.restype = C.c_int32
        L.orc_substep.argtypes = [P, C.c_double, C.c_double, C.c_double, dp]
        L.orc_substep.restype = C.c_int32
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_set_threads.restype = C.c_int32
        L.orc_refine.argtypes = [P, C.c_int32]
        L.orc_refine.restype = C.c_int32
        L.orc_regrid_sensors.argtypes = [P, dp, dp]
        L.orc_regrid_sensors.restype = C.c_int32
        L.orc_rebuild.argtypes = [P, C.c_int32]
        L.orc_rebuild.restype = C.c_int32
        L.orc_set_node_values.argtypes = [P, dp]
        L.orc_set_node_values.restype = C.c_int32
        L.orc_tree_import.argtypes = [P, C.c_int32, ip, ip, ip, ip, ip, ip, dp, dp, dp]
        L.orc_tree_import.restype = C.c_int32
        L.orc_rk3_range.argtypes = [P, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double]
        L.orc_rk3_range.restype = C.c_int32
        L.orc_neighbors_all.argtypes = [P, ip]
        L.orc_neighbors_all.restype = C.c_int32
        L.orc_stencil_all.argtypes = [P, ip, ip, dp, dp, ip, ip, ip, C.c_void_p]
        L.orc_stencil_all.restype = C.c_int32
        _lib = L
    return _lib


def make_params(**kw) -> OrcParams:
    """Defaults follow set_parameter (MAIN.F90:29-71)."""
    p = OrcParams()
    d = dict(xmin=0.0, xmax=10.0, ymin=0.0, ymax=10.0, h=0.5, gamma=GAMMA_REF, Ma=1.0,
             tstep=float(np.float32(1e-4)), cfl=float(np.float32(0.45)), t_end=10.0,
             iflag_localdt=0, iflag_amr=0, lvl_amr=0, pstep_adapt=30, nstep_adapt=0,
             refine_goal_rot=GAMMA_REF, refine_goal_div=GAMMA_REF,
             coarse_goal_rot=float(np.float32(0.4)), coarse_goal_div=float(np.float32(0.4)),
             amr_sensor=0, bc_xm=6, bc_xp=6, bc_ym=6, bc_yp=6, inflow=(0, 0, 0, 0),
             ic_kind=IC_VORTEX, seed=0, ic_param=(0,) * 8, pre_adapt=0, wls_local=0)
    d.update(kw)
    for k, v in d.items():
        if k in ("inflow", "ic_param"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = float(x)
        else:
            setattr(p, k, v)
    return p


class Oracle:
    """One reference-equivalent solver instance."""

    def __init__(self, **kw):
        self.L = lib()
        self.params = make_params(**kw)
        self.h = self.L.orc_new(C.byref(self.params))
        self.L.orc_init(self.h)

    def close(self):
        if self.h:
            self.L.orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- driving ----
    def step(self) -> int:
        return self.L.orc_step(self.h)

    def advance_rk3(self, dt, lam_x, lam_y):
        self.L.orc_advance_rk3(self.h, dt, lam_x, lam_y)

    def substep(self, dt, lam_x, lam_y):
        out = np.zeros((self.n_leaf, 5))
        self.L.orc_substep(self.h, dt, lam_x, lam_y, out.reshape(-1))
        return out

    # ---- accessors ----
    @property
    def n_leaf(self):
        return self.L.orc_n_leaf(self.h)

    @property
    def n_bc(self):
        return self.L.orc_n_bc(self.h)

    def dims(self):
        m, n = C.c_int32(), C.c_int32()
        self.L.orc_base_dims(self.h, C.byref(m), C.byref(n))
        return m.value, n.value

    def leaves(self):
        N = self.n_leaf
        x = np.zeros(N); y = np.zeros(N)
        lvl = np.zeros(N, np.int32); node = np.zeros(N, np.int32)
        prim = np.zeros((N, 5))
        self.L.orc_get_leaves(self.h, x, y, lvl, node, prim.reshape(-1))
        return dict(x=x, y=y, lvl=lvl, node=node, prim=prim)

    def bc_leaves(self):
        G = self.n_bc
        node = np.zeros(G, np.int32); prim = np.zeros((G, 5))
        self.L.orc_get_bc_leaves(self.h, node, prim.reshape(-1))
        return dict(node=node, prim=prim)

    def set_leaves(self, prim):
        self.L.orc_set_leaves(self.h, np.ascontiguousarray(prim, np.float64).reshape(-1))

    def scalars(self):
        out = np.zeros(8)
        self.L.orc_get_scalars(self.h, out)
        return dict(tt=out[0], tstep=out[1], lam_x=out[2], lam_y=out[3], l_max=int(out[4]),
                    it=int(out[5]))

    def residual(self):
        a = np.zeros(4); b = np.zeros(4)
        self.L.orc_get_residual(self.h, a, b)
        return a, b

    def tree(self):
        nn = self.L.orc_n_node(self.h)
        t = dict(father=np.zeros(nn, np.int32), son=np.zeros(nn, np.int32),
                 lvl=np.zeros(nn, np.int32), sp=np.zeros(nn, np.int32),
                 num=np.zeros((nn, 2), np.int32), cross=np.zeros(nn, np.int32),
                 x=np.zeros(nn), y=np.zeros(nn), prim7=np.zeros((nn, 7)))
        self.L.orc_tree_export(self.h, t["father"], t["son"], t["lvl"], t["sp"],
                               t["num"].reshape(-1), t["cross"], t["x"], t["y"],
                               t["prim7"].reshape(-1))
        return t

    def stencil_info(self, i, k):
        kind, node, m1, m2 = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        px, py = C.c_double(), C.c_double()
        nb = np.full(16, -1, np.int32)
        self.L.orc_stencil_info(self.h, i, k, C.byref(kind), C.byref(node), C.byref(px),
                                C.byref(py), C.byref(m1), C.byref(m2), nb)
        return dict(kind=kind.value, node=node.value, px=px.value, py=py.value, m1=m1.value,
                    m2=m2.value, nb=nb[: m1.value + m2.value].copy())

    def stencil_value(self, i, k):
        v = np.zeros(7)
        self.L.orc_stencil_value(self.h, i, k, v)
        return v

    def neighbors(self, node):
        out = np.zeros(8, np.int32)
        self.L.orc_neighbors(self.h, int(node), out)
        return out

    def neighbors_all(self):
        out = np.zeros((self.L.orc_n_node(self.h), 8), np.int32)
        self.L.orc_neighbors_all(self.h, out.reshape(-1))
        return out

    def stencil_all(self, values=False):
        """Classification (and values) of every interior leaf's 8 stencil entries."""
        N = self.n_leaf
        r = dict(kind=np.zeros((N, 8), np.int32), node=np.zeros((N, 8), np.int32), px=np.zeros((N, 8)),
                 py=np.zeros((N, 8)), m1=np.zeros((N, 8), np.int32), m2=np.zeros((N, 8), np.int32),
                 nb=np.zeros((N, 8, 9), np.int32))
        var = np.zeros((N, 8, 7)) if values else None
        self.L.orc_stencil_all(self.h, r["kind"].reshape(-1), r["node"].reshape(-1), r["px"].reshape(-1),
                               r["py"].reshape(-1), r["m1"].reshape(-1), r["m2"].reshape(-1),
                               r["nb"].reshape(-1), var.ctypes.data if values else None)
        if values:
            r["var"] = var
        return r

    def refine(self, node):
        assert self.L.orc_refine(self.h, int(node)) == 0, node

    def rebuild(self, reinit_ic=True):
        self.L.orc_rebuild(self.h, 1 if reinit_ic else 0)

    def set_node_values(self, prim7):
        self.L.orc_set_node_values(self.h, np.ascontiguousarray(prim7, np.float64).reshape(-1))

    def import_tree(self, t):
        rc = self.L.orc_tree_import(self.h, len(t["father"]), *[np.ascontiguousarray(t[k], np.int32).reshape(-1)
                                                                 for k in ("father", "son", "lvl", "sp", "num", "cross")],
                                    *[np.ascontiguousarray(t[k], np.float64).reshape(-1) for k in ("x", "y", "prim7")])
        assert rc == 0

    def rk3_range(self, i0, i1, dt, lam_x, lam_y):
        assert self.L.orc_rk3_range(self.h, i0, i1, dt, lam_x, lam_y) == 0

    def regrid_sensors(self, rot, div):
        """amr_regrid with the interior leaves' sensors given (cell_NS order)."""
        rot = np.ascontiguousarray(rot, np.float64)
        div = np.ascontiguousarray(div, np.float64)
        assert rot.shape == div.shape == (self.n_leaf,)
        assert self.L.orc_regrid_sensors(self.h, rot, div) == 0

    def sensors(self):
        N = self.n_leaf
        rot = np.zeros(N); div = np.zeros(N); s = np.zeros(2)
        ne = self.L.orc_sensors(self.h, rot, div, s)
        return rot, div, s, ne


# ---- pure kernel restatements ----
def wcns_z3(u1, u2, u3, R):
    out = np.zeros(4)
    lib().orc_wcns_z3(np.ascontiguousarray(u1, float), np.ascontiguousarray(u2, float),
                      np.ascontiguousarray(u3, float),
                      np.ascontiguousarray(np.asarray(R, float).T.reshape(-1)), out)
    return out


def eigen(direction, Ucons, gamma=GAMMA_REF, cv=None):
    if cv is None:
        cv = 1.0 / (gamma * (gamma - 1.0))
    R = np.zeros(16); Ri = np.zeros(16)
    lib().orc_eigen(direction, np.ascontiguousarray(Ucons, float), gamma, cv, R, Ri)
    return R.reshape(4, 4).T.copy(), Ri.reshape(4, 4).T.copy()


def flux_derivative(direction, prim5x5, d, lam, gamma=GAMMA_REF):
    out = np.zeros(4)
    cv = 1.0 / (gamma * (gamma - 1.0))
    lib().orc_flux_derivative(direction, np.ascontiguousarray(prim5x5, float).reshape(-1), d,
                              lam, gamma, cv, out)
    return out


def brinv(a):
    a = np.asarray(a, float)
    n = a.shape[0]
    r = np.zeros(n * n)
    flag = lib().orc_brinv(n, np.ascontiguousarray(a.T.reshape(-1)), r)
    return flag, r.reshape(n, n).T.copy()


def get_direction(xx, yy, x, y):
    return lib().orc_get_direction(xx, yy, x, y)


def wls_weno(pts, m1, xx, yy, cx, cy, lvl, h, x_factor=1.0):
    pts = np.ascontiguousarray(pts, float)
    out = np.zeros(4)
    lib().orc_wls_weno(pts.shape[0], m1, pts.reshape(-1), xx, yy, cx, cy, lvl, h, x_factor, out)
    return out


def smoothness(order, coeff, cx, cy, lvl, h, x_factor=1.0):
    return lib().orc_smoothness(order, np.ascontiguousarray(coeff, float), cx, cy, lvl, h, x_factor)


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's leaf loops (n <= 0: query)."""
    return lib().orc_set_threads(n)


def splitmix64(state: int, count: int):
    s = C.c_uint64(state)
    return [lib().orc_splitmix64(C.byref(s)) for _ in range(count)]
