"""Pins for the CPU oracle (oracle/fdm_oracle.c).

The reference has no tests, fixtures or golden vectors, and it does not compile with the
image's only Fortran compiler, so every oracle function is pinned here by something the
mathematics fixes: closed forms, invariants, brute force on small trees and convergence
orders. Each check is chosen so that a plausible slip (a dropped term, a wrong sign or index,
a transposed operand) fails it.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

G = O.GAMMA_REF


# ---------------------------------------------------------------- dense inverse (brinv)
def test_brinv_inverse_and_singular_flag():
    rng = np.random.default_rng(0)
    for n in (3, 6):
        for _ in range(20):
            a = rng.normal(size=(n, n)) + n * np.eye(n)
            flag, r = O.brinv(a)
            assert flag == 1
            assert np.allclose(a @ r, np.eye(n), atol=1e-13)
            assert np.allclose(r, np.linalg.inv(a), rtol=1e-12, atol=1e-13)
    # full pivoting must cope with a zero leading entry
    a = np.array([[0.0, 2.0, 1.0], [1.0, 0.0, 3.0], [4.0, 1.0, 0.0]])
    assert np.allclose(O.brinv(a)[1], np.linalg.inv(a), atol=1e-14)
    # exactly singular: flag 0 (XuanGua.F90:616-619)
    assert O.brinv(np.ones((3, 3)))[0] == 0


# ---------------------------------------------------------------- direction binning
def test_get_direction_octants():
    # dir 1 = east, counter-clockwise in 45-degree sectors centred on the axes
    expect = {(1, 0): 1, (1, 1): 2, (0, 1): 3, (-1, 1): 4, (-1, 0): 5, (-1, -1): 6, (0, -1): 7,
              (1, -1): 8}
    for (dx, dy), d in expect.items():
        assert O.get_direction(5.0 + dx, 3.0 + dy, 5.0, 3.0) == d
    # sector edges at 22.5 + 45k degrees
    for k in range(8):
        for off in (-1e-3, 1e-3):
            ang = math.radians(22.5 + 45 * k + off)
            d = O.get_direction(math.cos(ang), math.sin(ang), 0.0, 0.0)
            centre = (k + (1 if off > 0 else 0)) % 8 + 1
            assert d == centre


# ---------------------------------------------------------------- WCNS_Z3
def test_wcns_z3_closed_forms():
    # The linear weights are REAL(4) 1/11 and 10/11 widened to REAL(8) (nnd_center_RK.F90:331);
    # they sum to 1 + 2.98e-8, so smooth data is reproduced up to the factor 1/(d1+d2).
    d1, d2 = float(np.float32(1.0) / np.float32(11.0)), float(np.float32(10.0) / np.float32(11.0))
    assert d1 + d2 - 1.0 == pytest.approx(2.98e-8, rel=1e-2)
    I = np.eye(4)
    # constant data (binary fractions, so IS2 rounds to exactly 0; with epsilon = 1e-40 any
    # round-off in IS2 would flip the weights to the first-order candidate)
    c = np.array([1.25, -0.75, 2.0, 5.5])
    assert np.allclose(O.wcns_z3(c, c, c, I), c / (d1 + d2), rtol=1e-15, atol=0)
    # ... and that flip is the reference's behaviour: 1.3^2 sums leave IS2 ~ 1e-16 > 0
    c2 = np.array([1.3, 1.3, 1.3, 1.3])
    assert O.wcns_z3(c2, c2, c2, I)[0] == 1.3
    u1, u2, u3 = np.array([0.0, 1, 2, -3]), np.array([1.0, 2, 2.5, -1]), np.array([2.0, 3, 3, 1])
    lin = (-u1 + 6 * u2 + 3 * u3) / 8 / (d1 + d2)                                # linear data
    assert np.allclose(O.wcns_z3(u1, u2, u3, I), lin, rtol=1e-14, atol=1e-14)
    # a step between u2 and u3: IS1 = 0, the smooth side wins -> first order u2
    s1, s2, s3 = np.full(4, 1.0), np.full(4, 1.0), np.full(4, 3.0)
    assert np.allclose(O.wcns_z3(s1, s2, s3, I), s2, atol=1e-10)
    # back-projection hp = R uL uses R rows (not columns)
    R = np.arange(16.0).reshape(4, 4) / 7.0
    assert np.allclose(O.wcns_z3(c, c, c, R), R @ (c / (d1 + d2)), rtol=1e-14)


# ---------------------------------------------------------------- eigenvectors
def _jacobian(direction, U, gamma):
    def flux(U):
        r, mx, my, E = U
        u, v = mx / r, my / r
        p = (gamma - 1) * (E - 0.5 * r * (u * u + v * v))
        if direction == 0:
            return np.array([mx, mx * u + p, mx * v, u * (E + p)])
        return np.array([my, my * u, my * v + p, v * (E + p)])
    J = np.zeros((4, 4))
    for k in range(4):
        e = np.zeros(4); e[k] = 1e-6 * max(1.0, abs(U[k]))
        J[:, k] = (flux(U + e) - flux(U - e)) / (2 * e[k])
    return J


@pytest.mark.parametrize("direction", [0, 1])
def test_eigenvectors_diagonalise_flux_jacobian(direction):
    r, u, v, p = 1.2, 0.4, -0.3, 0.9
    E = p / (G - 1) + 0.5 * r * (u * u + v * v)
    U = np.array([r, r * u, r * v, E])
    R, Ri = O.eigen(direction, U)
    assert np.allclose(R @ Ri, np.eye(4), atol=1e-13)
    a = math.sqrt(G * p / r)
    vn = u if direction == 0 else v
    D = Ri @ _jacobian(direction, U, G) @ R
    assert np.allclose(D, np.diag([vn - a, vn, vn + a, vn]), atol=1e-6)


# ---------------------------------------------------------------- flux derivative
@pytest.mark.parametrize("direction", [0, 1])
def test_free_stream_preserved_exactly(direction):
    q = np.array([1.3, 0.6, -0.2, 0.0, 0.8])
    col = np.tile(q, (5, 1))
    out = O.flux_derivative(direction, col, 0.05, 2.0)
    assert np.all(out == 0.0)


@pytest.mark.parametrize("direction", [0, 1])
def test_flux_derivative_of_linear_density(direction):
    # rho linear along the column, u,v,p constant: the Euler residual of the mass equation is
    # u_n d(rho)/dn exactly for a 3rd-order scheme on linear data (no dissipation remains)
    d = 0.01
    rho = 1.0 + 0.1 * np.arange(-2, 3) * d
    col = np.zeros((5, 5))
    col[:, 0] = rho
    col[:, 1] = 0.7
    col[:, 2] = 0.4
    col[:, 4] = 1.0
    out = O.flux_derivative(direction, col, d, 3.0)
    un = 0.7 if direction == 0 else 0.4
    # 4/3 (F+ - F-)/d - 1/6 (E_{i+1} - E_{i-1})/d on linear data is the exact derivative
    # u_n d(rho)/dn, up to the reference's two REAL(4) artefacts: faces scaled by 1/(d1+d2)
    # and, in characteristic fields that are constant up to round-off (~1e-17), weights that
    # flip between linear and first order (epsilon = 1e-40), which leaves ~3e-8 |W| / d of
    # noise in dE. A sign slip or a dropped term is off by >= 1/3 of the value.
    u, v = 0.7, 0.4
    ke = 0.5 * (u * u + v * v)
    want = 0.1 * (np.array([u, u * u, u * v, u * ke]) if direction == 0
                  else np.array([v, u * v, v * v, v * ke]))
    assert want[0] == pytest.approx(un * 0.1)
    # noise bound: 4/3 * 2 * 3e-8 * max|W| / d with |W| <= 3 here
    assert np.allclose(out, want, rtol=0, atol=2.5e-5)


# ---------------------------------------------------------------- smoothness integrals
def _quad(f, x0, x1, y0, y1, n=40):
    xs, wx = np.polynomial.legendre.leggauss(n)
    X = 0.5 * (x1 - x0) * xs + 0.5 * (x1 + x0)
    Y = 0.5 * (y1 - y0) * xs + 0.5 * (y1 + y0)
    W = np.outer(wx, wx) * 0.25 * (x1 - x0) * (y1 - y0)
    XX, YY = np.meshgrid(X, Y, indexing="ij")
    return float(np.sum(W * f(XX, YY)))


def test_smoothness_indicator_integrals():
    # compute_smoothness integrates (k1 + k2 x + k3 y)^2 terms over the cell box
    # (XuanGua.F90:325-355) with k = (a1, 2 a3, a4) and (a2, a4, 2 a5); order 1 is (b1^2+b2^2)|box|
    a = np.array([0.3, 1.1, -0.7, 0.4, 0.25, -0.6])
    cx, cy, lvl, h = 2.3, 1.7, 1, 0.2
    hx = h / 2 ** (lvl + 1)
    x0, x1, y0, y1 = cx - hx, cx + hx, cy - hx, cy + hx
    area = (x1 - x0) * (y1 - y0)
    want = (_quad(lambda x, y: (a[1] + 2 * a[3] * x + a[4] * y) ** 2, x0, x1, y0, y1)
            + _quad(lambda x, y: (a[2] + a[4] * x + 2 * a[5] * y) ** 2, x0, x1, y0, y1)
            + ((2 * a[3]) ** 2 + a[4] ** 2 + (2 * a[5]) ** 2) * area)
    assert O.smoothness(2, a, cx, cy, lvl, h) == pytest.approx(want, rel=1e-11)
    b = np.array([0.5, -2.0, 0.3])
    assert O.smoothness(1, b, cx, cy, lvl, h) == pytest.approx((4.0 + 0.09) * area, rel=1e-13)


# ---------------------------------------------------------------- WLS-WENO
def _wls_points(xc=3.0, yc=2.0, hh=0.1, m1=4):
    axis = [(1, 0), (-1, 0), (0, 1), (0, -1), (0, 0)][:m1]
    diag = [(1, 1), (-1, 1), (-1, -1), (1, -1)]
    pts = [(xc + hh * i, yc + hh * j) for i, j in axis + diag]
    return np.array(pts), m1


def _lin(x, y, k):
    return [1.0 + 0.3 * x - 0.2 * y, 0.5 - x + 2 * y, -0.1 * x, 2.0 + 0.05 * y][k]


def _wls_linear_error(m1, xc, yc, hh):
    P, m1 = _wls_points(xc, yc, hh, m1)
    xx, yy = xc + hh / 2, yc
    pts = np.array([[x, y] + [_lin(x, y, k) for k in range(4)] for x, y in P])
    out = O.wls_weno(pts, m1, xx, yy, xc, yc, 1, 2 * hh)
    return np.max(np.abs(out - [_lin(xx, yy, k) for k in range(4)]))


@pytest.mark.parametrize("m1", [4, 5])
def test_wls_weno_reproduces_linear_and_constant_data(m1):
    # All three fits reproduce linear data, so the WENO combination does for any weights;
    # a sign slip in the combination (XuanGua.F90:304-306) breaks this.
    assert _wls_linear_error(m1, 0.0, 0.0, 0.1) < 1e-14
    P, m1 = _wls_points(0.0, 0.0, 0.1, m1)
    pts = np.array([[x, y, 1.5, -0.25, 0.75, 2.0] for x, y in P])
    out = O.wls_weno(pts, m1, 0.05, 0.0, 0.0, 0.0, 1, 0.2)
    assert np.allclose(out, [1.5, -0.25, 0.75, 2.0], rtol=1e-14)


def test_wls_weno_absolute_frame_conditioning():
    # The reference fits in absolute coordinates (XuanGua.F90:224-229): the normal equations
    # lose accuracy like (x/h)^4 away from the origin. The same fit about the cell centre
    # (the library's optional local frame) is exact again.
    e_origin = _wls_linear_error(4, 0.0, 0.0, 0.01)
    e_far = _wls_linear_error(4, 3.0, 2.0, 0.01)
    assert e_origin < 1e-13
    assert e_far > 1e-7
    # local frame: coordinates (x - xc)/h about the cell, box [-1/2, 1/2]^2
    P, m1 = _wls_points(3.0, 2.0, 0.01, 4)
    pts = np.array([[(x - 3.0) / 0.01, (y - 2.0) / 0.01] + [_lin(x, y, k) for k in range(4)] for x, y in P])
    out = O.wls_weno(pts, m1, 0.5, 0.0, 0.0, 0.0, 0, 1.0)
    assert np.max(np.abs(out - [_lin(3.005, 2.0, k) for k in range(4)])) < 1e-13


# ---------------------------------------------------------------- mesh, neighbours, leaf order
def _locate(t, m, n, h, xmin, ymin, x, y, lvl_max):
    """Brute force: deepest node with lvl <= lvl_max containing point (x, y)."""
    j = int(math.floor((x - xmin) / h)) + 1
    i = int(math.floor((y - ymin) / h)) + 1
    if i < 1 or i > n or j < 1 or j > m:
        return -1
    c = (i - 1) * m + (j - 1)
    while t["son"][c] >= 0 and t["lvl"][c] < lvl_max:
        s = t["son"][c]
        right = x > t["x"][c]
        up = y > t["y"][c]
        c = s + (2 if (right and up) else 1 if right else 3 if up else 0)
    return c


def _amr_oracle():
    o = O.Oracle(h=10.0 / 32, iflag_amr=1, lvl_amr=2, iflag_localdt=1, cfl=0.5, pstep_adapt=2)
    for _ in range(5):
        assert o.step() == 0
    return o


def test_face_neighbours_brute_force():
    o = _amr_oracle()
    t = o.tree()
    m, n = o.dims()
    h = 10.0 / 32
    xmin = ymin = -3 * h
    offs = {0: (0, 1), 1: (0, -1), 2: (1, 0), 3: (-1, 0)}  # up, down, right, left
    checked = 0
    for c in range(len(t["father"])):
        d = h / 2 ** t["lvl"][c]
        nb = o.neighbors(c)
        for k, (dx, dy) in offs.items():
            want = _locate(t, m, n, h, xmin, ymin, t["x"][c] + dx * d, t["y"][c] + dy * d, t["lvl"][c])
            assert nb[k] == want, (c, k)
            checked += 1
    assert checked > 10000


def test_leaf_order_is_base_rowmajor_then_morton():
    o = _amr_oracle()
    t = o.tree()
    L = o.leaves()

    def key(c):
        path = []
        while t["father"][c] >= 0:
            path.append({1: 0, 2: 1, 4: 2, 3: 3}[int(t["sp"][c])])
            c = t["father"][c]
        return (c, tuple(reversed(path)))

    keys = [key(c) for c in L["node"]]
    assert keys == sorted(keys)
    # every interior leaf (cross >= 0, no sons) is listed once
    inner = [c for c in range(len(t["father"])) if t["son"][c] < 0 and t["cross"][c] >= 0]
    assert sorted(inner) == sorted(L["node"].tolist())


def test_periodic_ghosts_copy_partner():
    o = O.Oracle(h=0.5)
    t = o.tree()
    m, n = o.dims()
    B = o.bc_leaves()
    for node, prim in zip(B["node"], B["prim"]):
        i, j = t["num"][node]
        ii = i + (n - 6) if i < 4 else i - (n - 6) if i > n - 3 else i
        jj = j + (m - 6) if j < 4 else j - (m - 6) if j > m - 3 else j
        assert np.array_equal(prim, t["prim7"][(ii - 1) * m + (jj - 1)][:5])


# ---------------------------------------------------------------- initial conditions
def test_vortex_initial_condition_closed_form():
    o = O.Oracle(h=0.5)
    L = o.leaves()
    x, y = L["x"], L["y"]
    rr = (x - 5) ** 2 + (y - 5) ** 2
    dT = -(G - 1) * 25 * np.exp(1 - rr) / (8 * G * np.pi ** 2)
    assert np.allclose(L["prim"][:, 0], (1 + dT) ** (1 / (G - 1)), rtol=1e-14)
    assert np.allclose(L["prim"][:, 1], 1 - (y - 5) * 5 / (2 * np.pi) * np.exp(0.5 * (1 - rr)), rtol=1e-14)
    assert np.allclose(L["prim"][:, 2], 1 + (x - 5) * 5 / (2 * np.pi) * np.exp(0.5 * (1 - rr)), rtol=1e-14)
    assert np.allclose(L["prim"][:, 4], (1 + dT) ** (G / (G - 1)), rtol=1e-14)


@pytest.mark.parametrize("mach", [2.0, 3.0])
def test_shock_state_satisfies_rankine_hugoniot(mach):
    g = 1.4
    o = O.Oracle(xmin=0, xmax=2, ymin=0, ymax=1, h=1 / 32, gamma=g, ic_kind=O.IC_SHOCK_BUBBLES,
                 seed=3, ic_param=(0.5, mach, 0, 0, 0, 0, 0), bc_xm=O.BC_INFLOW,
                 bc_xp=O.BC_TRANSMISSIVE, bc_ym=O.BC_TRANSMISSIVE, bc_yp=O.BC_TRANSMISSIVE)
    L = o.leaves()
    post = L["prim"][L["x"] < 0.5][0]
    r1, p1, u1 = 1.0, 1.0, 0.0
    r2, u2, p2 = post[0], post[1], post[4]
    s = mach * math.sqrt(g * p1 / r1)
    # fluxes in the shock frame are continuous
    assert r1 * (u1 - s) == pytest.approx(r2 * (u2 - s), rel=1e-13)
    assert r1 * (u1 - s) ** 2 + p1 == pytest.approx(r2 * (u2 - s) ** 2 + p2, rel=1e-13)
    e1 = p1 / ((g - 1) * r1) + 0.5 * (u1 - s) ** 2 + p1 / r1
    e2 = p2 / ((g - 1) * r2) + 0.5 * (u2 - s) ** 2 + p2 / r2
    assert e1 == pytest.approx(e2, rel=1e-13)


# ---------------------------------------------------------------- whole solver
def test_lambda_is_max_wavespeed_of_state():
    o = O.Oracle(h=0.5)
    p = o.leaves()["prim"]
    o.step()
    sc = o.scalars()
    a = np.sqrt(G * p[:, 4] / p[:, 0])
    assert sc["lam_x"] == np.max(np.abs(p[:, 1]) + a)
    assert sc["lam_y"] == np.max(np.abs(p[:, 2]) + a)


def _density_wave_error(n, t_end, cfl=0.4):
    o = O.Oracle(xmin=0, xmax=2, ymin=0, ymax=2, h=2.0 / n, ic_kind=O.IC_DENSITY_WAVE,
                 iflag_localdt=1, cfl=cfl, t_end=t_end)
    while o.scalars()["tt"] < t_end - 1e-14:
        assert o.step() == 0
    L = o.leaves()
    exact = 1 + 0.2 * np.sin(np.pi * (L["x"] + L["y"] - 2 * t_end))
    return float(np.mean(np.abs(L["prim"][:, 0] - exact)))


def test_density_wave_third_order_in_space():
    e = [_density_wave_error(n, 0.25) for n in (16, 32, 64)]
    orders = np.log2(np.array(e[:-1]) / np.array(e[1:]))
    assert orders[-1] > 2.6, (e, orders)


def test_rk3_is_shu_osher_composition_of_euler_substeps():
    # Time_advance_RK = Shu-Osher SSP-RK3 built from forward-Euler substeps E, blended on the
    # primitive variables (solution.F90:386-404): U1 = E(Un); U2 = 3/4 Un + 1/4 E(U1);
    # U3 = 1/3 Un + 2/3 E(U2). Composing the substeps by hand must give the same state bit
    # for bit (a wrong coefficient, or a blend on conserved variables, does not).
    kw = dict(h=10.0 / 32, iflag_amr=1, lvl_amr=2, pstep_adapt=2, iflag_localdt=1, cfl=0.5)
    o = O.Oracle(**kw)
    for _ in range(3):
        assert o.step() == 0  # gets level jumps into the mesh
    dt, lx, ly = 0.01, 3.1, 3.0
    un = o.leaves()["prim"].copy()
    s1 = o.substep(dt, lx, ly)
    u1 = 0.0 * un + 1.0 * s1
    o.set_leaves(u1)
    s2 = o.substep(dt, lx, ly)
    u2 = 3.0 / 4.0 * un + 1.0 / 4.0 * s2
    o.set_leaves(u2)
    s3 = o.substep(dt, lx, ly)
    u3 = 1.0 / 3.0 * un + 2.0 / 3.0 * s3
    o.set_leaves(un)
    o.advance_rk3(dt, lx, ly)
    assert np.array_equal(o.leaves()["prim"], u3)
    # the blend is on primitives: a conserved-variable blend differs
    m_cons = (1.0 / 3.0 * un[:, 0] * un[:, 1] + 2.0 / 3.0 * s3[:, 0] * s3[:, 1]) / u3[:, 0]
    assert np.max(np.abs(m_cons - u3[:, 1])) > 1e-12
