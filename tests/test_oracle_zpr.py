"""The zero-pressure right edge of sim_horiz_flow (fluid_2d::zpr) in the
oracle, pinned without the reference's solver library:

- the constant-density operators against their closed-form eigenvectors:
  MAC cells with q = 0 on the right edge face (mgs_mac.hh:50, a_cc = -3
  xxsp in the last column) have the eigenvectors cos(theta (i + 1/2)),
  theta = (k + 1/2) pi / m; the bilinear FEM stiffness with node column
  m-1 fixed (pfix, mgs_fem.cc:21) has the eigenvectors cos(theta i),
  theta = (k + 1/2) pi / (m - 1), with the boundary-halved weights of a
  tensor-product stiffness;
- the varying-density operators against matrices assembled from their
  definitions (tests/operators.py), and the solves against sparse direct
  solves;
- a horizontal-flow step: the inflow ghost columns, the fixed pressure
  column and the divergence of the corrected faces."""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spl

from tests import operators as ops
from tests import oracle_lib as ol
from tests.test_oracle_mg import fem_apply, fem_solve, mac_apply, mac_solve


@pytest.mark.parametrize("y_prd", [True, False])
@pytest.mark.parametrize("k,l", [(0, 1), (3, 2), (7, 5)])
def test_mac_zpr_eigenvectors(y_prd, k, l):
    m, n, dx, dy = 32, 24, 0.1, 0.07
    th = (k + 0.5) * math.pi / m
    ph = 2 * math.pi * l / n if y_prd else l * math.pi / n
    i, j = np.meshgrid(np.arange(m), np.arange(n))
    q = np.cos(th * (i + 0.5)) * (np.cos(ph * j) if y_prd else np.cos(ph * (j + 0.5)))
    lam = -4 * math.sin(th / 2) ** 2 / dx**2 - 4 * math.sin(ph / 2) ** 2 / dy**2
    y = mac_apply(m, n, dx, dy, None, q.ravel(), py=y_prd, zpr=True).reshape(n, m)
    assert np.abs(y - lam * q).max() <= 1e-11 * abs(lam)


@pytest.mark.parametrize("y_prd", [True, False])
@pytest.mark.parametrize("k,l", [(0, 1), (2, 3), (6, 4)])
def test_fem_zpr_eigenvectors(y_prd, k, l):
    m, n, dx, dy = 33, 24 if y_prd else 25, 0.1, 0.07   # m nodes, the last one fixed
    th = (k + 0.5) * math.pi / (m - 1)
    ph = 2 * math.pi * l / n if y_prd else l * math.pi / (n - 1)
    qx, qy = np.cos(th * np.arange(m)), np.cos(ph * np.arange(n))
    qx[-1] = 0.0   # cos((k + 1/2) pi), exactly
    # 1-D stiffness and mass eigenvalues of linear elements
    lkx, lmx = (2 - 2 * math.cos(th)) / dx, (4 + 2 * math.cos(th)) * dx / 6
    lky, lmy = (2 - 2 * math.cos(ph)) / dy, (4 + 2 * math.cos(ph)) * dy / 6
    wx = np.ones(m); wx[0] = 0.5
    wy = np.ones(n)
    if not y_prd:
        wy[0] = wy[-1] = 0.5
    lam = lkx * lmy + lmx * lky
    q = np.outer(qy, qx)
    y = fem_apply(m, n, dx, dy, None, q.ravel(), py=y_prd, zpr=True).reshape(n, m)
    want = lam * np.outer(wy * qy, wx * qx)
    assert np.abs(y - want).max() <= 1e-12 * lam


def _density(m, n, seed):
    R = ops.blob_density(m, n, ratio=3.0, seed=seed)
    return R


@pytest.mark.parametrize("y_prd", [True, False])
def test_mac_zpr_operator_and_solve(y_prd):
    m, n, dx, dy = 32, 16, 0.125, 0.1
    irho = ops.mac_irho_from_rho(_density(m, n, 1), m, n)
    A = ops.mac_matrix(m, n, dx, dy, irho, py=y_prd, zpr=True)
    x = np.random.default_rng(0).standard_normal(m * n)
    y = mac_apply(m, n, dx, dy, irho, x, py=y_prd, zpr=True)
    assert np.abs(y - A @ x).max() <= 1e-12 * np.abs(A @ x).max()
    b = np.random.default_rng(1).standard_normal(m * n)
    z, cyc, rr = mac_solve(m, n, dx, dy, irho, b, tol=1e-12, py=y_prd, zpr=True)
    assert cyc > 0 and rr <= 1e-12
    zd = spl.spsolve(A.tocsc(), b)   # definite: no mean removed
    assert np.abs(z - zd).max() <= 1e-9 * np.abs(zd).max()


@pytest.mark.parametrize("y_prd", [True, False])
def test_fem_zpr_operator_and_solve(y_prd):
    m, n = 33, (16 if y_prd else 17)
    dx, dy = 0.125, 0.1
    R = _density(m, n, 2)
    if y_prd:
        irho = (1.0 / R[0:n, 0:m]).ravel()
    else:
        irho = ops.fem_irho_from_rho(R, m - 1, n - 1)
    A = ops.fem_matrix(m, n, dx, dy, irho, py=y_prd, zpr=True)
    x = np.random.default_rng(3).standard_normal(m * n)
    y = fem_apply(m, n, dx, dy, irho, x, py=y_prd, zpr=True)
    assert np.abs(y - A @ x).max() <= 1e-12 * np.abs(A @ x).max()
    b = np.random.default_rng(4).standard_normal(m * n)
    z, cyc, rr = fem_solve(m, n, dx, dy, irho, b, tol=1e-12, py=y_prd, zpr=True)
    bb = b.copy().reshape(n, m)
    bb[:, -1] = 0.0     # mgs_fem.cc:85
    zd = spl.spsolve(A.tocsc(), bb.ravel())
    assert cyc > 0 and rr <= 1e-12
    assert np.abs(z.reshape(n, m)[:, -1]).max() == 0.0
    assert np.abs(z - zd).max() <= 1e-9 * np.abs(zd).max()


def test_horiz_flow_step_properties():
    spec = ol.horiz_flow_spec(64, 32, U=1.0)
    o = ol.OracleFluid(spec, mg_tol=1e-12)
    h = 4.0 / 64
    for k in range(3):
        o.step(0.2 * h)
        u, v, p = o.get("u"), o.get("v"), o.get("p")
        assert np.all(u[:, 0:2] == 1.0) and np.all(v[:, 0:2] == 0.0)   # sim_horiz_flow::boundary
        assert np.all(p[2:-2, 2 + 64] == 0.0)                         # node column m fixed at zero
        # right ghost columns: field::extrap
        assert np.array_equal(u[2:-2, -2], 2 * u[2:-2, -3] - u[2:-2, -4])
    o.close()


def test_uniform_flow_stays_uniform():
    """Closed form: with no solid, y periodic, the uniform inflow u = U,
    v = 0 is a steady state of the horizontal-flow step, and it is kept
    exactly only if every boundary piece agrees with it: the inflow ghosts
    (U, 0), the inflow face left as cell 0's own state (not a zeroed wall
    face) in the MAC source, the ghost velocity in the FEM source left of
    node 0 (not zero), the fixed node column and no mean removed."""
    spec = ol.horiz_flow_spec(32, 16, U=1.0)
    spec["objects"] = []
    o = ol.OracleFluid(spec, mg_tol=1e-12)
    for k in range(4):
        o.step(0.02)
    u, v, p = o.get("u"), o.get("v"), o.get("p")
    assert np.all(u[2:-2, 2:-2] == 1.0) and np.all(v[2:-2, 2:-2] == 0.0) and np.all(p[2:-2, 2:-2] == 0.0)
    o.close()
