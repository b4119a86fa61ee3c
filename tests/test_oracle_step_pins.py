"""The oracle's step sweeps (oracle/fluid_oracle.cc) pinned phase by phase
by closed forms.

fluid_2d.cc cannot be compiled here (it needs the absent TGMG), so each
sweep of the restatement is held to what the continuum operator it
discretises gives exactly on fields where the discretisation has no
truncation error (linear and bilinear data), and the whole step to the
second-order convergence of an exact Navier-Stokes solution. Each check is
derived from the mathematics, not from the oracle's expressions, so a
dropped term, a wrong sign, a swapped component or a wrong factor in a
sweep fails it (tests/test_oracle_pin_mutants.py plants such slips in a
copy of the oracle and shows that these tests catch them).

  stress         fluid_2d.cc:381-600   viscous stress of a linear velocity;
                 neo-Hookean stress of an affine reference map, with and
                 without actuation, plus the solid's extra viscosity
  ustar          fluid_2d.cc:1278-1291 u* = u - dt (u.grad) u + dt div(s)/rho
                 on linear face fields and stresses
  fem_source     fluid_2d.cc:1297-1309 -(dx dy/dt) div u* at interior nodes
                 for bilinear u*; the sum over all nodes is zero with walls
  copy_pressure  fluid_2d.cc:1317-1348 a linear solution minus its exact mean
  update_velocity fluid_2d.cc:1354-1374 u = u* - dt grad(p)/rho for linear p
  whole step     fluid_2d.cc:298-369   a decaying shear wave carried by a
                 uniform flow in a periodic box converges at second order
"""
import math

import numpy as np
import pytest

from tests import oracle_lib as ol


def fluid_only(m, visc=0.01, x_prd=1, y_prd=1, n=None):
    return dict(m=m, n=n or m, ax=-1.0, bx=1.0, ay=-1.0, by=1.0, visc=visc, rhof=1.0, tw=2.5, pinw=2.5,
                w_scale=3.0, w_force=20.0, ev_mult=0.8, dt_pad=0.4, dt_ev_pad=0.8,
                x_prd=x_prd, y_prd=y_prd, vel=0, objects=[])


def grids(f, m, n=None):
    """Cell-centre coordinates (x, y) over the ghosted (n+4) x (m+4) grid of
    the [-1,1]^2 box, the node coordinates of the cell slots (node (i,j) is
    cell (i,j)'s lower left corner), and the spacings."""
    n = n or m
    dx, dy = 2.0 / m, 2.0 / n
    xc, yc = np.meshgrid(-1 + (np.arange(-2, m + 2) + 0.5) * dx, -1 + (np.arange(-2, n + 2) + 0.5) * dy)
    return xc, yc, xc - 0.5 * dx, yc - 0.5 * dy, dx, dy


def test_fluid_stress_of_linear_velocity():
    # sigma = mu (grad u + grad u^T) for u = a x + b y + c, v = d x - a y + e
    m, n, mu = 24, 14, 0.013
    f = ol.OracleFluid(fluid_only(m, visc=mu, n=n))
    xc, yc, _, _, dx, dy = grids(f, m, n)
    a, b, c, d, e = 0.7, -1.3, 0.2, 0.45, -0.1
    f.set("u", a * xc + b * yc + c)
    f.set("v", d * xc - a * yc + e)
    f.phase("stress")
    sl = (slice(2, n + 3), slice(2, m + 3))   # every left and lower face the sweep writes
    for name, want in (("s11", 2 * mu * a), ("s21", mu * (b + d)),
                       ("s12", mu * (b + d)), ("s22", -2 * mu * a)):
        got = f.get(name)[sl]
        assert np.abs(got - want).max() <= 1e-12 * abs(mu), name


def _full_solid(m, G, rhos, kind=ol.FULL, p9=None, n=None):
    spec = fluid_only(m, n=n)
    spec["objects"] = [(kind, p9 or [0.0] * 9, G, rhos, 1.0, 0)]
    spec["solid_prd_bc"] = int(kind == ol.FULL)
    return spec


@pytest.mark.parametrize("with_velocity", [False, True])
def test_solid_stress_of_affine_reference_map(with_velocity):
    # deep inside a solid (phi < -eps: tf = 1, no transition term) the
    # stress of the reference map X = J x + x0 is the neo-Hookean
    # G (B - (tr B + 1)/3 I) with B = F F^T and F = adj(J) = det(J) J^-1
    # (the inverse deformation gradient), plus the extra viscosity times
    # the plain velocity gradient (fluid_2d.cc:510-560)
    m, n, G, rhos = 24, 18, 1.7, 2.0
    f = ol.OracleFluid(_full_solid(m, G, rhos, n=n))
    xc, yc, _, _, dx, dy = grids(f, m, n)
    J = np.array([[1.1, 0.25], [-0.15, 0.8]])
    f.set("X", J[0, 0] * xc + J[0, 1] * yc + 0.3, 0)
    f.set("Y", J[1, 0] * xc + J[1, 1] * yc - 0.2, 0)
    f.set("phi", -np.ones_like(xc), 0)
    gu = np.array([[0.6, -0.35], [0.9, -0.6]]) if with_velocity else np.zeros((2, 2))
    f.set("u", gu[0, 0] * xc + gu[0, 1] * yc + 0.1)
    f.set("v", gu[1, 0] * xc + gu[1, 1] * yc)
    f.phase("stress")
    F = np.linalg.det(J) * np.linalg.inv(J)
    B = F @ F.T
    sig = G * (B - (np.trace(B) + 1) / 3 * np.eye(2))
    ev = 0.8 * max(dx, dy) * rhos * math.sqrt(G / rhos)   # obj_field.cc:35-49 with ev_mult 0.8
    sl = (slice(2, n + 3), slice(2, m + 3))
    want = {"s11": sig[0, 0] + ev * gu[0, 0], "s21": sig[1, 0] + ev * gu[1, 0],
            "s12": sig[0, 1] + ev * gu[0, 1], "s22": sig[1, 1] + ev * gu[1, 1]}
    for name, w in want.items():
        assert np.abs(f.get(name)[sl] - w).max() <= 1e-12 * G, name


def test_solid_stress_with_actuation():
    # obj_flapper deep inside its actuated rod: the stress is the
    # neo-Hookean one of the actuated F (actuation pinned separately in
    # tests/test_oracle_actuated.py), evaluated at each face's mid reference
    # point, i.e. the first column of F scaled by s = exp(con Y_face), the
    # second by 1/s
    m, G, rhos, t = 32, 1.0, 1.0, 2.0
    flap = [0.0, 0.0, 0.05, 0.4, 0.3, 0.4, 0.0, 2.2, 8.0]   # cx cy cr cl ar al theta mstretch T
    f = ol.OracleFluid(_full_solid(m, G, rhos, ol.FLAPPER, flap))
    f.set_time(t)
    xc, yc, _, _, dx, dy = grids(f, m)
    J = np.array([[1.05, 0.1], [-0.05, 0.97]])
    # a reference map that keeps every face inside the actuated region
    X = 0.05 * (J[0, 0] * xc + J[0, 1] * yc)
    Y = 0.05 * (J[1, 0] * xc + J[1, 1] * yc)
    f.set("X", X, 0)
    f.set("Y", Y, 0)
    f.set("phi", -np.ones_like(xc), 0)
    f.set("u", 0 * xc)
    f.set("v", 0 * xc)
    f.phase("stress")
    con = -math.log(flap[7]) / flap[4] * math.sin(2 * math.pi * t / flap[8]) ** 12
    Jh = 0.05 * J
    F0 = np.linalg.det(Jh) * np.linalg.inv(Jh)
    for face, names in (("left", ("s11", "s21")), ("down", ("s12", "s22"))):
        # mid reference Y of the left face of cell (i,j) (or the lower face)
        Yf = 0.5 * (Y + (np.roll(Y, 1, axis=1) if face == "left" else np.roll(Y, 1, axis=0)))
        s = np.exp(con * Yf)
        Fa, Fb, Fc, Fd = F0[0, 0] * s, F0[0, 1] / s, F0[1, 0] * s, F0[1, 1] / s
        sa, sb, sd = Fa * Fa + Fb * Fb, Fa * Fc + Fb * Fd, Fc * Fc + Fd * Fd
        tr = (sa + sd + 1) / 3
        w = (G * (sa - tr), G * sb) if face == "left" else (G * sb, G * (sd - tr))
        sl = (slice(3, m + 2), slice(3, m + 2))
        for name, want in zip(names, w):
            got = f.get(name)[sl]
            assert np.abs(got - want[sl]).max() <= 1e-11 * np.abs(want[sl]).max(), (face, name)


def test_ustar_on_linear_fields():
    # compute_ustar with linear face velocities L and linear stresses:
    # u* = u - dt (L_ul(c) dL_ul/dx + L_vd(c) dL_ud/dy) + dt (ds11/dx + ds12/dy)/rho
    # (the conservative face differences are exact on linear data;
    # fluid_2d.cc:1278-1291, acceleration without pressure :1019-1039)
    m, n, dt = 16, 10, 0.013
    f = ol.OracleFluid(fluid_only(m, n=n))
    xc, yc, xl, yd, dx, dy = grids(f, m, n)
    lin = lambda a, b, c, x, y: a + b * x + c * y
    P = dict(ul=(0.3, 0.5, -0.2), vd=(-0.1, 0.25, 0.4), ud=(0.2, -0.3, 0.6), vl=(0.05, 0.7, -0.45),
             s11=(0.1, 0.9, -0.3), s12=(-0.2, 0.4, 1.1), s21=(0.3, -0.8, 0.2), s22=(0.0, 0.35, -0.6))
    u = 0.3 * np.sin(3 * xc) + yc
    v = np.cos(2 * yc) - 0.1 * xc
    rho = 1.5 + 0.5 * np.sin(xc + 2 * yc)
    f.set("rho", rho)
    f.set("u", u)
    f.set("v", v)
    f.set("ul", lin(*P["ul"], xl, yc))
    f.set("vl", lin(*P["vl"], xl, yc))
    f.set("ud", lin(*P["ud"], xc, yd))
    f.set("vd", lin(*P["vd"], xc, yd))
    for k in ("s11", "s21"):
        f.set(k, lin(*P[k], xl, yc))
    for k in ("s12", "s22"):
        f.set(k, lin(*P[k], xc, yd))
    f.phase("ustar", dt)
    Lul = lin(*P["ul"], xc, yc)
    Lvd = lin(*P["vd"], xc, yc)
    ustar = u - dt * (Lul * P["ul"][1] + Lvd * P["ud"][2]) + dt * (P["s11"][1] + P["s12"][2]) / rho
    vstar = v - dt * (Lul * P["vl"][1] + Lvd * P["vd"][2]) + dt * (P["s21"][1] + P["s22"][2]) / rho
    sl = (slice(2, n + 2), slice(2, m + 2))
    assert np.abs(f.get("c0")[sl] - ustar[sl]).max() <= 1e-13
    assert np.abs(f.get("c1")[sl] - vstar[sl]).max() <= 1e-13


def test_fem_source_of_bilinear_velocity():
    # nodal source = -(dx dy / dt) div u* at interior nodes for bilinear
    # u* (fluid_2d.cc:1297-1309); with no-slip walls the ghost u* is zero
    # and the sources of all nodes sum to zero (fluid_2d.cc:839-872)
    m, n, dt = 20, 12, 0.03
    f = ol.OracleFluid(fluid_only(m, x_prd=0, y_prd=0, n=n))
    xc, yc, xn, yn, dx, dy = grids(f, m, n)
    a = (0.2, 0.7, -0.4, 0.9)
    b = (-0.3, 0.15, 0.55, -0.6)
    f.set("c0", a[0] + a[1] * xc + a[2] * yc + a[3] * xc * yc)
    f.set("c1", b[0] + b[1] * xc + b[2] * yc + b[3] * xc * yc)
    src = f.phase("fem_source", dt)
    assert src.shape == (n + 1, m + 1)
    X, Y = xn[2:n + 3, 2:m + 3], yn[2:n + 3, 2:m + 3]
    div = (a[1] + a[3] * Y) + (b[2] + b[3] * X)
    want = -(dx * dy / dt) * div
    assert np.abs(src[1:-1, 1:-1] - want[1:-1, 1:-1]).max() <= 1e-12 * np.abs(want).max()
    assert abs(src.sum()) <= 1e-12 * np.abs(src).sum()


@pytest.mark.parametrize("prd", [0, 1])
def test_copy_pressure_of_linear_solution(prd):
    # p = z - (trapezoid mean of z); for a linear z on the walled box the
    # trapezoid rule is exact and the mean is z at the centre
    # (fluid_2d.cc:1317-1348). A periodic solution must be periodic: a
    # constant, which copy_pressure removes entirely.
    m, n = 16, 10
    f = ol.OracleFluid(fluid_only(m, x_prd=prd, y_prd=prd, n=n))
    mf, nf = (m, n) if prd else (m + 1, n + 1)
    _, _, xn, yn, dx, dy = grids(f, m, n)
    X, Y = xn[2:2 + nf, 2:2 + mf], yn[2:2 + nf, 2:2 + mf]
    z = (0.0 * X + 1.75) if prd else (0.6 * X - 1.1 * Y + 2.5)
    f.phase("copy_pressure", 0.0, z)
    p = f.get("p")[2:n + 3, 2:m + 3]
    want = 0.0 * p if prd else 0.6 * xn[2:n + 3, 2:m + 3] - 1.1 * yn[2:n + 3, 2:m + 3]
    assert np.abs(p - want).max() <= 1e-13


def test_update_velocity_of_linear_pressure():
    # u = u* - dt grad(p)/rho with the cell gradient of the four corner
    # nodes: exact for linear p (fluid_2d.cc:1354-1374)
    m, n, dt = 16, 22, 0.021
    f = ol.OracleFluid(fluid_only(m, n=n))
    xc, yc, xn, yn, dx, dy = grids(f, m, n)
    rng = np.random.default_rng(5)
    rho = rng.uniform(1.0, 3.0, xc.shape)
    c0, c1 = rng.standard_normal(xc.shape), rng.standard_normal(xc.shape)
    al, be = 0.8, -1.4
    f.set("p", al * xn + be * yn + 0.3)
    f.set("rho", rho)
    f.set("c0", c0)
    f.set("c1", c1)
    f.phase("update_velocity", dt)
    sl = (slice(2, n + 2), slice(2, m + 2))
    assert np.abs(f.get("u")[sl] - (c0 - dt * al / rho)[sl]).max() <= 1e-13
    assert np.abs(f.get("v")[sl] - (c1 - dt * be / rho)[sl]).max() <= 1e-13


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("U,V", [(0.4, 0.3), (-0.4, -0.3), (0.4, -0.3)])
def test_predictor_on_travelling_shear(axis, U, V):
    # u = U + b (y - V t), v = V (axis 0; axis 1 exchanges the roles) is an
    # exact Navier-Stokes solution with constant pressure and stress. The
    # predictor's values are exact on it (fluid_2d.cc:877-1121, 1278-1291):
    #   - first-pass edge values: the data at the face point traced back half
    #     a step along the normal velocity;
    #   - half-time edge values: the solution at the face and t + dt/2;
    #   - u*: the solution at t + dt.
    # Periodic along the flow's variation, walls across it: cells four or
    # more away from the walls are checked.
    m, n, dt, b = 20, 24, 0.02, 0.7
    f = ol.OracleFluid(fluid_only(m, x_prd=1 - axis, y_prd=axis, n=n))
    xc, yc, _, _, dx, dy = grids(f, m, n)
    if axis == 0:
        ex = lambda x, y, t: (U + b * (y - V * t), V + 0 * x)
    else:
        ex = lambda x, y, t: (U + 0 * x, V + b * (x - U * t))
    u0, v0 = ex(xc, yc, 0.0)
    f.set("u", u0)
    f.set("v", v0)
    f.phase("stress")
    f.phase("normal_derivatives", dt)
    sl = (slice(6, n - 2), slice(6, m - 2))
    faces = (("l", -0.5, 0), ("r", 0.5, 0), ("d", 0, -0.5), ("u", 0, 0.5))
    for c, ox, oy in faces:
        # normal velocity of the face direction traces the point back
        tx = ox * dx - (0.5 * dt * (U if ox else 0) if ox else 0)
        ty = oy * dy - (0.5 * dt * (V if oy else 0) if oy else 0)
        uu, vv = ex(xc + tx, yc + ty, 0.0)
        if axis == 0:   # u varies along y only; v = V
            uu, vv = ex(xc, yc + ty, 0.0)
        else:
            uu, vv = ex(xc + tx, yc, 0.0)
        assert np.abs(f.get("u" + c)[sl] - uu[sl]).max() <= 1e-14, ("u" + c, "first pass")
        assert np.abs(f.get("v" + c)[sl] - vv[sl]).max() <= 1e-14, ("v" + c, "first pass")
    for ph in ("godunov_tang", "edge_bc", "tangential", "half_time_edges"):
        f.phase(ph, dt)
    for c, ox, oy in faces:
        uu, vv = ex(xc + ox * dx, yc + oy * dy, 0.5 * dt)
        assert np.abs(f.get("u" + c)[sl] - uu[sl]).max() <= 1e-14, ("u" + c, "half time")
        assert np.abs(f.get("v" + c)[sl] - vv[sl]).max() <= 1e-14, ("v" + c, "half time")
    # (the MAC projection is left out: with walls across the flow the
    # solution is not the walled box's, and its face divergence is zero in
    # the periodic direction anyway)
    for ph in ("godunov", "edge_bc", "stress"):
        f.phase(ph, dt)
    f.phase("ustar", dt)
    u1, v1 = ex(xc, yc, dt)
    assert np.abs(f.get("c0")[sl] - u1[sl]).max() <= 1e-14
    assert np.abs(f.get("c1")[sl] - v1[sl]).max() <= 1e-14


@pytest.mark.parametrize("varying", [False, True])
def test_mac_projection_removes_a_gradient(varying):
    # faces that are the discrete gradient (times the face inverse density)
    # of a cell field q0 with no flow through the walls are projected to zero
    # (fluid_2d.cc:1175-1236): the solve recovers q0 up to a constant and the
    # correction subtracts the same gradient
    m, n = 16, 12
    spec = fluid_only(m, x_prd=0, y_prd=0, n=n)
    if varying:   # a solid far outside the box makes the density vary
        spec["objects"] = [(ol.CIRCLE, [5.0, 5.0, 0.1, 0.0, 0.0, 0.0], 1.0, 3.0, 1.0, 0)]
    f = ol.OracleFluid(spec, mg_tol=1e-13)
    xc, yc, _, _, dx, dy = grids(f, m, n)
    rng = np.random.default_rng(11)
    rho = rng.uniform(1.0, 3.0, xc.shape) if varying else np.ones_like(xc)
    f.set("rho", rho)
    q0 = rng.standard_normal(xc.shape)
    el = 0.5 / rho + 0.5 / np.roll(rho, 1, axis=1)
    ed = 0.5 / rho + 0.5 / np.roll(rho, 1, axis=0)
    ul = el * (q0 - np.roll(q0, 1, axis=1)) / dx
    vd = ed * (q0 - np.roll(q0, 1, axis=0)) / dy
    ul[:, :3] = 0; ul[:, m + 2:] = 0     # no flow through the walls (faces i = 0 and m)
    vd[:3, :] = 0; vd[n + 2:, :] = 0
    f.set("ul", ul)
    f.set("vd", vd)
    f.phase("mac_projection")
    sl = (slice(2, n + 2), slice(2, m + 2))
    scale = max(np.abs(ul).max(), np.abs(vd).max())
    assert np.abs(f.get("ul")[sl]).max() <= 1e-10 * scale
    assert np.abs(f.get("vd")[sl]).max() <= 1e-10 * scale


def test_reference_map_advection_of_affine_map():
    # an affine reference map X = J x + x0 carried by a uniform flow (U,V)
    # is X(x,t) = J (x - (U,V) t) + x0 exactly: limited slopes are exact on
    # linear data, the half-time edge values are exact at (face, t + dt/2),
    # and the conservative update is exact (fluid_2d.cc:915-950, 1082-1108,
    # 1112-1121, 1242-1273). Deep inside a solid (status 0) update_reference_map
    # sets nX to the new map and X to the half-time map. The ghost layers
    # are not refreshed between phases here, so only cells three or more
    # away from the edges are checked.
    m, n, dt, U, V = 24, 20, 0.01, 0.45, -0.3
    f = ol.OracleFluid(_full_solid(m, 1.0, 1.0, n=n))
    xc, yc, _, _, dx, dy = grids(f, m, n)
    J = np.array([[1.2, -0.35], [0.25, 0.9]])
    x0 = (0.1, -0.2)
    aff = lambda x, y, k: J[k, 0] * x + J[k, 1] * y + x0[k]
    f.set("X", aff(xc, yc, 0), 0)
    f.set("Y", aff(xc, yc, 1), 0)
    f.set("phi", -np.ones_like(xc), 0)
    f.set("s", np.zeros_like(xc), 0)
    f.set("u", U + 0 * xc)
    f.set("v", V + 0 * xc)
    f.phase("stress")
    for ph in ("normal_derivatives", "godunov_tang", "edge_bc", "tangential", "half_time_edges"):
        f.phase(ph, dt)
    sl = (slice(5, n - 1), slice(5, m - 1))
    # half-time edge values: the map at the face point traced back half a
    # step along the flow
    for nm, k, ox, oy in (("Xl", 0, -0.5, 0), ("Xr", 0, 0.5, 0), ("Xd", 0, 0, -0.5), ("Xu", 0, 0, 0.5),
                          ("Yl", 1, -0.5, 0), ("Yr", 1, 0.5, 0), ("Yd", 1, 0, -0.5), ("Yu", 1, 0, 0.5)):
        want = aff(xc + ox * dx - 0.5 * U * dt, yc + oy * dy - 0.5 * V * dt, k)
        assert np.abs(f.get(nm, 0)[sl] - want[sl]).max() <= 1e-14, nm
    for ph in ("godunov", "edge_bc", "mac_projection", "update_reference_map"):
        f.phase(ph, dt)
    for k, (a, b) in enumerate((("X", "nX"), ("Y", "nY"))):
        half = aff(xc - 0.5 * U * dt, yc - 0.5 * V * dt, k)
        new = aff(xc - U * dt, yc - V * dt, k)
        assert np.abs(f.get(b, 0)[sl] - new[sl]).max() <= 1e-14, b
        assert np.abs(f.get(a, 0)[sl] - half[sl]).max() <= 1e-14, a


def test_translating_full_solid_both_directions():
    # full steps of obj_full with solid periodic boundaries under the uniform
    # flow (U,V): a rigid translation X = x - U t, Y = y - V t, velocity
    # unchanged (extends the x-only translation of
    # tests/test_oracle_solid_periodic.py to both directions)
    spec = ol.conv_spec(32, 2, 280)
    spec["vel"] = 0
    f = ol.OracleFluid(spec, mg_tol=1e-12)
    U, V, dt = 0.4, -0.3, 0.01
    f.set("u", U + 0 * f.get("u"))
    f.set("v", V + 0 * f.get("v"))
    X0, Y0 = f.get("X").copy(), f.get("Y").copy()
    for _ in range(5):
        f.step(dt)
    sl = (slice(2, -2), slice(2, -2))
    assert np.abs(f.get("u")[sl] - U).max() <= 1e-14 and np.abs(f.get("v")[sl] - V).max() <= 1e-14
    assert np.abs(f.get("X")[sl] - (X0[sl] - 5 * U * dt)).max() <= 1e-14
    assert np.abs(f.get("Y")[sl] - (Y0[sl] - 5 * V * dt)).max() <= 1e-14


def shear_wave(x, y, t, U, V, A, nu):
    """Exact Navier-Stokes solution in the periodic [-1,1]^2 box: a shear
    wave along k = pi (1,1) with its velocity along (1,-1)/sqrt 2, carried by
    the uniform flow (U,V) and decaying as exp(-nu |k|^2 t); its pressure is
    constant ((u.grad) u = 0 for the wave part)."""
    k2 = 2 * math.pi ** 2
    s = A * math.exp(-nu * k2 * t) * np.sin(math.pi * ((x - U * t) + (y - V * t))) / math.sqrt(2)
    return U + s, V - s


def run_wave(m, U, V, A, nu, T):
    f = ol.OracleFluid(fluid_only(m, visc=nu), mg_tol=1e-12)
    xc, yc, _, _, h, _ = grids(f, m)
    u0, v0 = shear_wave(xc, yc, 0.0, U, V, A, nu)
    f.set("u", u0)
    f.set("v", v0)
    f.phase("stress")
    nsteps = int(round(T / (0.25 * h)))
    dt = T / nsteps
    for _ in range(nsteps):
        f.step(dt)
    ue, ve = shear_wave(xc, yc, T, U, V, A, nu)
    sl = (slice(2, m + 2), slice(2, m + 2))
    err = np.mean(np.abs(f.get("u")[sl] - ue[sl]) + np.abs(f.get("v")[sl] - ve[sl]))
    return err, np.abs(f.get("p")[sl]).max()


@pytest.mark.parametrize("U,V", [(0.5, 0.25), (0.1, -0.05)])
def test_shear_wave_converges_at_second_order(U, V):
    # the full step (normal slopes, Godunov upwinding, transverse terms,
    # half-time edges with viscous acceleration, MAC projection, stress,
    # u*, FEM projection, pressure update) on an exact solution, mean
    # absolute error at t = 0.4 with dt = h/4 (the slope limiter clips the
    # wave's extrema, which costs order in the maximum norm but not in the
    # mean); the second flow reverses direction across the wave, so every
    # branch of the upwinding is taken
    A, nu, T = 0.3, 0.001, 0.4
    runs = [run_wave(m, U, V, A, nu, T) for m in (24, 48, 96)]
    e = [r[0] for r in runs]
    o1, o2 = math.log2(e[0] / e[1]), math.log2(e[1] / e[2])
    assert 0.5 * (o1 + o2) > 1.8 and min(o1, o2) > 1.4, (e, o1, o2)
    assert e[2] < 1e-3 * A
    # the exact pressure is constant: the computed one is truncation-sized
    assert runs[2][1] < 1e-3 * A * A
