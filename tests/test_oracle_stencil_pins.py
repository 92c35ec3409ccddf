"""Pins for the oracle's reference-specific index logic: stencil selection at level jumps.

Each check compares the oracle (oracle/fdm_oracle.c) with an independent model written from
the reference's definitions, not from the oracle's code:

* diagonal neighbours (neighbor.F90:199-396, cached by idnei only for cross >= 0,
  neighbor.F90:421-431): the geometric neighbour (deepest node no finer than the cell that
  contains the cell centre moved one cell diagonally), nulled by a data table of the
  reference's asymmetric rules;
* the stencil gathers Tip..Tjmm (XuanGua.F90:401-590): which node each entry reads, its kind
  and virtual point from geometry, and on a linear field every value reproduced (WLS-WENO
  reproduces linear data exactly; interpA returns the centre value; the child fallback copy
  returns the coarse cell's own value);
* search_cell_5 / get_direction / add_to_neighbor_list (XuanGua.F90:11-43,152-177) and the
  parent / child stencils with their m < 8 widening and m < 7 fallbacks (XuanGua.F90:45-150):
  a brute-force depth-first search (son1..son4, the 4x repeat kept) with octants from exact
  comparisons against tan(22.5 deg), compared list by list; plus hand-built trees whose
  expected lists for m = 6, 7, 8 were derived by hand;
* the periodic ghost fill BC_cyc at levels >= 2 (solution.F90:109-168), whose son path is
  walked leaf-first (reversed);
* the 2:1 balance that gridmodify1/gridmodify4 leave behind (mesh.F90:305-395) and the rot_div
  sensors on linear and quadratic fields (solution.F90:558-580).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

TAN22 = math.sqrt(2.0) - 1.0  # tan(22.5 deg): the octant edges of get_direction (XuanGua.F90:40-42)
SP_QUADRANT = {1: (-1, -1), 2: (1, -1), 3: (1, 1), 4: (-1, 1)}  # son1..son4 = SW, SE, NE, NW (mesh.F90:204-248)

# Null rules of the four diagonal finders (neighbor.F90:199-396), as data: for each corner and
# son position, "sibling" (the father's other son, never null), "both" (null only when the
# horizontal or the vertical face neighbour is missing), "H" / "V" (null when that face
# neighbour is missing or coarser than the cell). Base cells: null when either is missing.
DIAG_RULES = {
    "upleft": {2: "sibling", 4: "both", 1: "H", 3: "V"},     # neighbor.F90:216,219,229,236
    "upright": {1: "sibling", 3: "both", 2: "H", 4: "V"},    # neighbor.F90:266,269,279,286
    "downright": {4: "sibling", 2: "both", 3: "H", 1: "V"},  # neighbor.F90:316,319,329,336
    "downleft": {3: "sibling", 1: "both", 4: "H", 2: "V"},   # neighbor.F90:366,369,379,386
}
# corner -> (sx, sy) and its column in orc_neighbors' output (up, down, right, left, upleft,
# upright, downleft, downright)
CORNERS = {"upleft": ((-1, 1), 4), "upright": ((1, 1), 5), "downleft": ((-1, -1), 6), "downright": ((1, -1), 7)}


class Model:
    """Geometry of an exported oracle tree (canonical numbering)."""

    def __init__(self, o, h, xmin, ymin):
        self.o = o
        self.t = o.tree()
        self.m, self.n = o.dims()
        self.h = h
        self.X0, self.Y0 = xmin - 3 * h, ymin - 3 * h
        t = self.t
        self.size = h / 2.0 ** t["lvl"]
        self.nodes = len(t["father"])

    def locate(self, x, y, lvl_max):
        """Deepest node with lvl <= lvl_max containing (x, y); -1 outside the base grid."""
        t = self.t
        j = int(math.floor((x - self.X0) / self.h)) + 1
        i = int(math.floor((y - self.Y0) / self.h)) + 1
        if i < 1 or i > self.n or j < 1 or j > self.m:
            return -1
        c = (i - 1) * self.m + (j - 1)
        while t["son"][c] >= 0 and t["lvl"][c] < lvl_max:
            right, up = x > t["x"][c], y > t["y"][c]
            c = t["son"][c] + (2 if (right and up) else 1 if right else 3 if up else 0)
        return c

    def offset(self, c, ax, ay):
        """Node at cell c's centre moved by (ax, ay) of its own spacing, no finer than c."""
        d = self.size[c]
        return self.locate(self.t["x"][c] + ax * d, self.t["y"][c] + ay * d, self.t["lvl"][c])

    def diagonal(self, c, corner):
        """Expected cached diagonal neighbour (idnei + the four finders)."""
        t = self.t
        if t["cross"][c] < 0:
            return -1  # idnei caches diagonals only for cross >= 0 (neighbor.F90:421-431)
        (sx, sy), _ = CORNERS[corner]
        H, V = self.offset(c, sx, 0), self.offset(c, 0, sy)
        geo = self.offset(c, sx, sy)
        if t["father"][c] < 0:
            return -1 if (H < 0 or V < 0) else geo
        rule = DIAG_RULES[corner][int(t["sp"][c])]
        if rule == "sibling":
            return geo
        if rule == "both":
            return -1 if (H < 0 or V < 0) else geo
        nb = H if rule == "H" else V
        return -1 if (nb < 0 or t["lvl"][nb] < t["lvl"][c]) else geo

    def neighbours(self, c):
        """The 8 cached neighbours in orc_neighbors' order."""
        out = [self.offset(c, 0, 1), self.offset(c, 0, -1), self.offset(c, 1, 0), self.offset(c, -1, 0)]
        out += [self.diagonal(c, k) for k in ("upleft", "upright", "downleft", "downright")]
        return out

    def leaves_dfs(self, c):
        """Leaves below c as search_cell_5 visits them: son1..son4, each subtree four times
        (XuanGua.F90:169-175)."""
        t = self.t
        if t["son"][c] < 0:
            yield c
            return
        for _ in range(4):
            for k in range(4):
                yield from self.leaves_dfs(t["son"][c] + k)


def octant(dx, dy):
    """get_direction (XuanGua.F90:33-43) by exact comparisons: 1 = east, counter-clockwise in
    45-degree sectors centred on the axes and the diagonals."""
    if dx > 0 and abs(dy) < TAN22 * dx:
        return 1
    if dy > 0 and abs(dx) < TAN22 * dy:
        return 3
    if dx < 0 and abs(dy) < TAN22 * -dx:
        return 5
    if dy < 0 and abs(dx) < TAN22 * -dy:
        return 7
    return {(1, 1): 2, (-1, 1): 4, (-1, -1): 6, (1, -1): 8}[(int(np.sign(dx)), int(np.sign(dy)))]


def brute_lists(M, c, parent):
    """parent_value_weno / child_value_weno stencil lists (XuanGua.F90:45-150) by brute force."""
    t = M.t
    x, y = t["x"][c], t["y"][c]
    used, l1, l2 = set(), [], []

    def search(start):
        for q in M.leaves_dfs(start):
            d = octant(t["x"][q] - x, t["y"][q] - y)
            if d in used:
                continue
            target = l1 if d in (1, 3, 5, 7) else l2
            if q not in target:
                target.append(q)
            used.add(d)

    nb = M.neighbours(c)  # up, down, right, left, upleft, upright, downleft, downright
    if parent:
        for k in range(4):
            search(t["son"][c] + k)
        if len(l1) + len(l2) < 8:
            for q in (nb[3], nb[2], nb[1], nb[0]):  # left, right, down, up
                if q >= 0:
                    search(q)
    else:
        l1.append(c)
        for q in (nb[6], nb[3], nb[4], nb[1], nb[0], nb[7], nb[2], nb[5]):  # DL, L, UL, D, U, DR, R, UR
            if q >= 0:
                search(q)
    return l1, l2


# ---------------------------------------------------------------- meshes from the reference's own AMR
def _vortex():
    o = O.Oracle(h=10.0 / 32, iflag_amr=1, lvl_amr=2, iflag_localdt=1, cfl=0.5, pstep_adapt=2, wls_local=1)
    for _ in range(5):
        assert o.step() == 0
    return o, 10.0 / 32, 0.0, 0.0


def _riemann():
    o = O.Oracle(xmin=0, xmax=1, ymin=0, ymax=1, h=1 / 24, gamma=1.4, bc_xm=1, bc_xp=1, bc_ym=1, bc_yp=1,
                 ic_kind=O.IC_RIEMANN4, seed=1, ic_param=(0.5, 0.5), iflag_localdt=1, cfl=0.5, t_end=0.2,
                 iflag_amr=1, lvl_amr=3, pstep_adapt=2, amr_sensor=1, wls_local=1)
    for _ in range(5):
        assert o.step() == 0
    return o, 1 / 24, 0.0, 0.0


def periodic_wave_kwargs():
    # periodic density wave, refinement to level 3 reaching the periodic ghosts (level 2 there)
    return dict(xmin=0, xmax=2, ymin=0, ymax=2, h=2 / 24, ic_kind=O.IC_DENSITY_WAVE, iflag_localdt=1, cfl=0.5,
                iflag_amr=1, lvl_amr=3, pstep_adapt=2, amr_sensor=1, wls_local=1, refine_goal_rot=1.0,
                refine_goal_div=1.0)


def _periodic():
    o = O.Oracle(**periodic_wave_kwargs())
    for _ in range(7):
        assert o.step() == 0
    return o, 2 / 24, 0.0, 0.0


MESHES = {"vortex": _vortex, "riemann": _riemann, "periodic": _periodic}


@pytest.fixture(scope="module", params=list(MESHES))
def mesh(request):
    o, h, xmin, ymin = MESHES[request.param]()
    return request.param, o, Model(o, h, xmin, ymin)


# ---------------------------------------------------------------- neighbours
def test_diagonal_neighbours_geometric_with_null_table(mesh):
    name, o, M = mesh
    nb = o.neighbors_all()
    nulls_by_rule = 0
    for c in range(M.nodes):
        want = M.neighbours(c)
        assert list(nb[c]) == want, (name, c)
        if M.t["cross"][c] >= 0 and M.t["father"][c] >= 0:
            for corner in ("upleft", "upright", "downleft", "downright"):
                (sx, sy), col = CORNERS[corner]
                if want[col] < 0 and M.offset(c, sx, sy) >= 0:
                    nulls_by_rule += 1
    # the asymmetric null rules are exercised (a geometric neighbour exists but is not cached)
    assert nulls_by_rule > 20, (name, nulls_by_rule)


def test_reference_amr_leaves_meshes_2to1_balanced(mesh):
    # gridmodify4 refines a face neighbour, and that neighbour's neighbour, when it is two or
    # more levels coarser (mesh.F90:355-395), lvlAMR passes after gridmodify1: no interior leaf
    # sees a leaf two levels coarser across a face or one cell beyond it
    name, o, M = mesh
    t = M.t
    checked = 0
    for c in np.flatnonzero((t["son"] < 0) & (t["cross"] >= 0)):
        L, d = t["lvl"][c], M.size[c]
        for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            q = M.offset(c, dx, dy)
            if q < 0:
                continue
            assert L - t["lvl"][q] <= 1, (name, c)
            s = M.size[q]
            q2 = M.locate(t["x"][c] + dx * (d + s), t["y"][c] + dy * (d + s), L)
            if q2 >= 0:
                assert L - t["lvl"][q2] <= 1, (name, c)
            checked += 1
    assert checked > 1000


# ---------------------------------------------------------------- stencil gathers
def _linear(x, y):
    return np.stack([1.0 + 0.3 * x - 0.2 * y, 0.5 - 0.1 * x + 0.2 * y, -0.1 * x + 0.05 * y,
                     np.zeros_like(x), 2.0 + 0.05 * y - 0.02 * x], axis=-1)


def test_stencil_gathers_geometry_kinds_and_linear_values(mesh):
    name, o, M = mesh
    t = M.t
    prim7 = np.zeros((M.nodes, 7))
    prim7[:, :5] = _linear(t["x"], t["y"])
    o.set_node_values(prim7)
    S = o.stencil_all(values=True)
    L = o.leaves()
    kinds = np.zeros(5, int)
    axis = [(1, 0), (-1, 0), (0, 1), (0, -1)]
    worst = 0.0
    for i, c in enumerate(L["node"]):
        d = M.size[c]
        for k in range(8):
            ax, ay = axis[k & 3]
            f = 2.0 if k >= 4 else 1.0
            px, py = t["x"][c] + f * ax * d, t["y"][c] + f * ay * d
            tg = M.locate(px, py, t["lvl"][c])
            kind = int(S["kind"][i, k])
            kinds[kind] += 1
            assert S["node"][i, k] == tg, (name, i, k)
            same = t["lvl"][tg] == t["lvl"][c]
            if same and t["son"][tg] < 0:
                assert kind == 0
            elif same:
                assert kind in (1, 2)
                assert (S["px"][i, k], S["py"][i, k]) == (t["x"][tg], t["y"][tg])
                assert abs(t["x"][tg] - px) < 1e-12 * (1 + abs(px)) and abs(t["y"][tg] - py) < 1e-12 * (1 + abs(py))
            else:
                assert kind in (3, 4)
                assert abs(S["px"][i, k] - px) <= 1e-15 * (1 + abs(px)) and abs(S["py"][i, k] - py) <= 1e-15 * (1 + abs(py))
                # the virtual point lies inside the coarse cell it is interpolated in
                half = M.size[tg] / 2
                assert abs(px - t["x"][tg]) < half and abs(py - t["y"][tg]) < half
            want = _linear(np.array(t["x"][tg]), np.array(t["y"][tg])) if kind == 4 else _linear(np.array(px), np.array(py))
            got = S["var"][i, k]
            err = float(np.max(np.abs(got[[0, 1, 2, 4]] - want[[0, 1, 2, 4]])))
            worst = max(worst, err)
            assert err < 1e-12, (name, i, k, kind, got, want)
    print(name, "kinds", kinds, "worst", worst)
    assert kinds[0] > 0 and kinds[1] > 0 and kinds[3] > 0
    if name != "vortex":
        assert kinds[4] > 0  # child fallback copies at ghost cells (all diagonals null: m = 5)


def test_stencil_gathers_reference_frame_linear():
    # the reference's absolute-coordinate frame (wls_local = 0, XuanGua.F90:224-229): linear data
    # is still reproduced, to the conditioning of the absolute-frame normal equations
    o = O.Oracle(h=10.0 / 32, iflag_amr=1, lvl_amr=2, iflag_localdt=1, cfl=0.5, pstep_adapt=2)
    for _ in range(5):
        assert o.step() == 0
    M = Model(o, 10.0 / 32, 0.0, 0.0)
    prim7 = np.zeros((M.nodes, 7))
    prim7[:, :5] = _linear(M.t["x"], M.t["y"])
    o.set_node_values(prim7)
    S = o.stencil_all(values=True)
    L = o.leaves()
    axis = [(1, 0), (-1, 0), (0, 1), (0, -1)]
    wls = 0
    for i, c in enumerate(L["node"]):
        d = M.size[c]
        for k in range(8):
            if S["kind"][i, k] not in (1, 3):
                continue
            wls += 1
            ax, ay = axis[k & 3]
            f = 2.0 if k >= 4 else 1.0
            want = _linear(np.array(M.t["x"][c] + f * ax * d), np.array(M.t["y"][c] + f * ay * d))
            assert np.max(np.abs(S["var"][i, k][[0, 1, 2, 4]] - want[[0, 1, 2, 4]])) < 1e-6
    assert wls > 500


def test_wls_stencil_lists_brute_force(mesh):
    # every parent / child stencil list equals the brute-force search (order included), and the
    # fallbacks are taken exactly when m < 7 (XuanGua.F90:87,143)
    name, o, M = mesh
    S = o.stencil_all()
    memo = {}
    m_seen = set()
    for i in range(S["kind"].shape[0]):
        for k in range(8):
            kind = int(S["kind"][i, k])
            if kind == 0:
                continue
            tg = int(S["node"][i, k])
            parent = kind in (1, 2)
            key = (tg, parent)
            if key not in memo:
                memo[key] = brute_lists(M, tg, parent)
            l1, l2 = memo[key]
            m = len(l1) + len(l2)
            m_seen.add((parent, m))
            assert (S["m1"][i, k], S["m2"][i, k]) == (len(l1), len(l2)), (name, i, k)
            assert list(S["nb"][i, k][:m]) == l1 + l2, (name, i, k)
            if parent:
                assert kind == (2 if m < 7 else 1)
            else:
                assert kind == (4 if m < 7 else 3)
    print(name, "distinct WLS cells", len(memo), "(parent, m) seen", sorted(m_seen))
    assert len(memo) > 50


# ---------------------------------------------------------------- hand-built trees, m = 6, 7, 8
def _hand_tree(refine_centres):
    """Oracle on [0,10]^2 with h = 1 (16 x 16 base cells with ghosts, centres at j - 3.5),
    then Refine_cell on the nodes with the given (x, y, lvl), in order."""
    o = O.Oracle(h=1.0)
    for x, y, lvl in refine_centres:
        o.refine(_node(o.tree(), x, y, lvl))
    o.rebuild(reinit_ic=True)
    return o


def _node(t, x, y, lvl):
    hit = np.flatnonzero((np.abs(t["x"] - x) < 1e-12) & (np.abs(t["y"] - y) < 1e-12) & (t["lvl"] == lvl))
    assert len(hit) == 1, (x, y, lvl)
    return int(hit[0])


def _entry(o, x, y, lvl, k):
    """Stencil entry k of the interior leaf centred at (x, y) on level lvl."""
    L = o.leaves()
    i = int(np.flatnonzero((np.abs(L["x"] - x) < 1e-12) & (np.abs(L["y"] - y) < 1e-12) & (L["lvl"] == lvl))[0])
    return o.stencil_info(i, k)


def _ids(t, pts):
    return [_node(t, x, y, l) for x, y, l in pts]


def test_hand_built_child_stencil_m7_and_singular_list2():
    # F = base cell centred (4.5, 4.5), refined; F.son4 (NW, centre (4.25, 4.75)) refined again.
    # The level-2 leaf q = F.son4.son1 at (4.125, 4.625) reads, below it (Tjm), the coarser
    # leaf c = F.son1 (4.25, 4.25): child_value_weno at (4.125, 4.375). c is a son1 whose left
    # (base (3.5, 4.5)) and down (base (4.5, 3.5)) neighbours are coarser, so upleft (rule H of
    # sp 1) and downright (rule V of sp 1) are null; downleft ("both") is base (3.5, 3.5) and
    # upright is the sibling F.son3. Search order DL, L, UL, D, U, DR, R, UR from c's centre:
    # DL octant 6 -> list 2; L (-0.75, 0.25) octant 5 -> list 1; D (0.25, -0.75) octant 7;
    # U = F.son4, refined: its first leaf son1 (-0.125, 0.375) is octant 3, the others repeat
    # it; R = F.son2 octant 1; UR = F.son3 octant 2. m = 5 + 2 = 7 -> WLS, with a 2-point
    # list 2 (rank-deficient 3x3 normal matrix: brinv's singular branch, XuanGua.F90:616-619).
    o = _hand_tree([(4.5, 4.5, 0), (4.25, 4.75, 1)])
    t = o.tree()
    e = _entry(o, 4.125, 4.625, 2, 3)
    assert e["kind"] == 3 and e["node"] == _node(t, 4.25, 4.25, 1)
    assert (e["px"], e["py"]) == (4.125, 4.375)
    l1 = _ids(t, [(4.25, 4.25, 1), (3.5, 4.5, 0), (4.5, 3.5, 0), (4.125, 4.625, 2), (4.75, 4.25, 1)])
    l2 = _ids(t, [(3.5, 3.5, 0), (4.75, 4.75, 1)])
    assert (e["m1"], e["m2"]) == (5, 2)
    assert list(e["nb"]) == l1 + l2
    # Tjmm of the same leaf: the -1y neighbour is coarser, so the same coarse cell at (4.125, 4.125)
    e2 = _entry(o, 4.125, 4.625, 2, 7)
    assert e2["kind"] == 3 and (e2["px"], e2["py"]) == (4.125, 4.125) and list(e2["nb"]) == l1 + l2


def test_hand_built_child_stencil_m8():
    # As above, with the base cell below F refined too: down(c) is now its son4 (4.25, 3.75) at
    # c's level, so downright (rule V) is that cell's right neighbour, son3 (4.75, 3.75), in
    # octant 8. Lists: list 1 = c, L, D = (4.25, 3.75), U's first leaf, R; list 2 = DL (base
    # (3.5, 3.5), octant 6), DR (octant 8), UR (octant 2). m = 8 -> WLS.
    o = _hand_tree([(4.5, 4.5, 0), (4.25, 4.75, 1), (4.5, 3.5, 0)])
    t = o.tree()
    e = _entry(o, 4.125, 4.625, 2, 3)
    assert e["kind"] == 3
    l1 = _ids(t, [(4.25, 4.25, 1), (3.5, 4.5, 0), (4.25, 3.75, 1), (4.125, 4.625, 2), (4.75, 4.25, 1)])
    l2 = _ids(t, [(3.5, 3.5, 0), (4.75, 3.75, 1), (4.75, 4.75, 1)])
    assert (e["m1"], e["m2"]) == (5, 3)
    assert list(e["nb"]) == l1 + l2


def test_hand_built_child_stencil_m6_falls_back_to_copy():
    # A ghost cell (cross < 0) has no cached diagonals, so its child stencil is itself plus
    # what its four face searches find. Base columns 1-3 (x = -2.5, -1.5, -0.5) are ghosts;
    # take the ghost g = (-0.5, 4.5) and refine its right neighbour, the interior base cell
    # (0.5, 4.5): the level-1 leaf (0.25, 4.25) reads g as its -1x value at (-0.25, 4.25).
    # g's up neighbour (-0.5, 5.5) is refined, and that cell's SE son (-0.25, 5.25) again.
    # From g's centre: L (-1.5, 4.5) octant 5, D (-0.5, 3.5) octant 7; the up search finds its
    # son1 (-0.75, 5.25) at (-0.25, 0.75), octant 3, first, then in the refined son2 the leaf
    # (-0.125, 5.125) at (0.375, 0.625), octant 2 -> list 2 (its other leaves repeat octants 2
    # and 3); R = the refined interior cell, whose first leaf (0.25, 4.25) at (0.75, -0.25) is
    # octant 1. m = 5 + 1 = 6 < 7 -> the fallback copies g's own value (XuanGua.F90:143-144).
    o = _hand_tree([(0.5, 4.5, 0), (-0.5, 5.5, 0), (-0.25, 5.25, 1)])
    t = o.tree()
    g = _node(t, -0.5, 4.5, 0)
    assert t["cross"][g] < 0
    e = _entry(o, 0.25, 4.25, 1, 1)
    assert e["kind"] == 4 and e["node"] == g
    assert (e["px"], e["py"]) == (-0.25, 4.25)
    l1 = _ids(t, [(-0.5, 4.5, 0), (-1.5, 4.5, 0), (-0.5, 3.5, 0), (-0.75, 5.25, 1), (0.25, 4.25, 1)])
    l2 = _ids(t, [(-0.125, 5.125, 2)])
    assert (e["m1"], e["m2"]) == (5, 1)
    assert list(e["nb"]) == l1 + l2
    # the copy: g's value, not an interpolant at the virtual point
    L = o.leaves()
    i = int(np.flatnonzero(L["node"] == _node(t, 0.25, 4.25, 1))[0])
    assert np.array_equal(o.stencil_value(i, 1), t["prim7"][g])


def test_hand_built_parent_stencils_m6_m7_m8():
    # F = base (4.5, 4.5) refined, F.son1 (4.25, 4.25) refined, its son1 p = (4.125, 4.125)
    # refined once more. The level-2 leaf (4.375, 4.125) reads p as its -1x value (same level,
    # refined): parent_value_weno(p). p's four sons are the diagonal octants 6, 8, 2, 4
    # (list 2, son order); m = 4 < 8 widens the search to L, R, D, U (XuanGua.F90:64-83):
    # L = base (3.5, 4.5), two levels coarser, at (-0.625, 0.375) from p: octant 4, used;
    # R = (4.375, 4.125): octant 1; D = base (4.5, 3.5) at (0.375, -0.625): octant 8, used;
    # U = (4.125, 4.375): octant 3. m = 2 + 4 = 6 < 7 -> interpA (XuanGua.F90:87-88).
    base = [(4.5, 4.5, 0), (4.25, 4.25, 1), (4.125, 4.125, 2)]
    o = _hand_tree(base)
    t = o.tree()
    p = _node(t, 4.125, 4.125, 2)
    sons = [(4.0625, 4.0625, 3), (4.1875, 4.0625, 3), (4.1875, 4.1875, 3), (4.0625, 4.1875, 3)]
    e = _entry(o, 4.375, 4.125, 2, 1)
    assert e["kind"] == 2 and e["node"] == p
    assert (e["m1"], e["m2"]) == (2, 4)
    assert list(e["nb"]) == _ids(t, [(4.375, 4.125, 2), (4.125, 4.375, 2)] + sons)
    # base (4.5, 3.5) refined: D is its son4 (4.25, 3.75) at (0.125, -0.375), octant 7 -> m = 7, WLS
    o = _hand_tree(base + [(4.5, 3.5, 0)])
    t = o.tree()
    e = _entry(o, 4.375, 4.125, 2, 1)
    assert e["kind"] == 1 and (e["m1"], e["m2"]) == (3, 4)
    assert list(e["nb"]) == _ids(t, [(4.375, 4.125, 2), (4.25, 3.75, 1), (4.125, 4.375, 2)] + sons)
    # base (3.5, 4.5) refined as well: L is its son2 (3.75, 4.25) at (-0.375, 0.125), octant 5 -> m = 8
    o = _hand_tree(base + [(4.5, 3.5, 0), (3.5, 4.5, 0)])
    t = o.tree()
    e = _entry(o, 4.375, 4.125, 2, 1)
    assert e["kind"] == 1 and (e["m1"], e["m2"]) == (4, 4)
    assert list(e["nb"]) == _ids(t, [(3.75, 4.25, 1), (4.375, 4.125, 2), (4.25, 3.75, 1), (4.125, 4.375, 2)] + sons)


# ---------------------------------------------------------------- periodic ghosts (BC_cyc)
def _interpA(t, c):
    """Restriction mean (XuanGua.F90:359-398): ((s1 + s2) + s3) + s4, times 1/4, recursively."""
    if t["son"][c] < 0:
        return t["prim7"][c][:5]
    s = [_interpA(t, t["son"][c] + k) for k in range(4)]
    return 0.25 * (((s[0] + s[1]) + s[2]) + s[3])


def test_periodic_ghosts_follow_reversed_son_path():
    # BC_cyc (solution.F90:109-168): the periodic partner base cell is the ghost's base cell
    # shifted by m - 6 columns or n - 6 rows; the partner is then descended by the ghost's son
    # positions taken leaf-first (sp(1) is the ghost's own position, solution.F90:142-159) while
    # it has sons, and the ghost takes interpA of the node reached.
    o, h, _, _ = _periodic()
    t = o.tree()
    m, n = o.dims()
    B = o.bc_leaves()
    lvl2 = differs = 0
    for g, prim in zip(B["node"], B["prim"]):
        i, j = t["num"][g]
        if j <= 3:
            j2, i2 = j + (m - 6), i
        elif j >= m - 2:
            j2, i2 = j - (m - 6), i
        elif i <= 3:
            i2, j2 = i + (n - 6), j
        else:
            i2, j2 = i - (n - 6), j
        path = []  # sp from the ghost leaf up to its base cell
        c = g
        while t["father"][c] >= 0:
            path.append(int(t["sp"][c]))
            c = t["father"][c]
        p = fwd = (i2 - 1) * m + (j2 - 1)
        for s in path:  # leaf-first
            if t["son"][p] < 0:
                break
            p = t["son"][p] + s - 1
        for s in reversed(path):  # root-first (the geometric image)
            if t["son"][fwd] < 0:
                break
            fwd = t["son"][fwd] + s - 1
        lvl2 += len(path) >= 2
        differs += p != fwd
        assert np.array_equal(prim, _interpA(t, p)), g
    print("ghosts", len(B["node"]), "at level >= 2:", lvl2, "partner differs from the geometric image:", differs)
    assert lvl2 > 50 and differs > 20


# ---------------------------------------------------------------- rot_div sensors
def test_sensors_closed_forms():
    # rot = |v_x - u_y| dy^1.5 and div = (|rho_x| + |rho_y|) dy^1.5 from central differences of
    # the +-1 stencil values (solution.F90:558-580): exact on a linear field at every leaf
    # (level jumps included, the stencil values reproduce it), and on a quadratic field at
    # uniform leaves (central differences are exact for quadratics)
    o, h, xmin, ymin = _vortex()
    M = Model(o, h, xmin, ymin)
    t = M.t
    prim7 = np.zeros((M.nodes, 7))
    prim7[:, :5] = _linear(t["x"], t["y"])
    o.set_node_values(prim7)
    rot, div, s2, ne = o.sensors()
    L = o.leaves()
    dy = h / 2.0 ** L["lvl"]
    # v_x - u_y = -0.1 - 0.2; rho_x = 0.3, rho_y = -0.2
    assert np.allclose(rot, 0.3 * dy ** 1.5, rtol=1e-10, atol=0)
    assert np.allclose(div, 0.5 * dy ** 1.5, rtol=1e-10, atol=0)
    assert ne == o.n_leaf and s2[0] == pytest.approx(np.sum(rot ** 2), rel=1e-12)
    # quadratic: u = x y, v = x^2 + y^2 / 2, rho = 1 + x^2 + x y  ->  v_x - u_y = 2x - x = x,
    # rho_x = 2x + y, rho_y = x
    u = O.Oracle(h=10.0 / 32)
    Mu = Model(u, 10.0 / 32, 0.0, 0.0)
    x, y = Mu.t["x"], Mu.t["y"]
    q = np.zeros((Mu.nodes, 7))
    q[:, 0], q[:, 1], q[:, 2], q[:, 4] = 1 + x * x + x * y, x * y, x * x + 0.5 * y * y, 1.0
    u.set_node_values(q)
    rot, div, _, _ = u.sensors()
    Lu = u.leaves()
    d15 = (10.0 / 32) ** 1.5
    assert np.allclose(rot, np.abs(Lu["x"]) * d15, rtol=1e-11, atol=1e-13)
    assert np.allclose(div, (np.abs(2 * Lu["x"] + Lu["y"]) + np.abs(Lu["x"])) * d15, rtol=1e-11, atol=1e-13)
