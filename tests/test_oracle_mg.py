"""Oracle projection solvers (oracle/mg_oracle.cc) pinned to mathematics.

The reference's multigrid library (TGMG) is not in the reference tree, so the
oracle solves the reference's systems with its own solver. What these tests
pin is that the oracle's operators ARE the discretisations the reference
defines (mgs_mac.hh:41-102, mgs_fem.hh:61-136), assembled here independently
from their definitions (tests/operators.py), and that its solves agree with
dense direct solves and converge at the discretisation's order.
"""
import ctypes

import numpy as np
import pytest

from tests import operators as ops
from tests import oracle_lib as ol

D, I = ctypes.c_double, ctypes.c_int


def _lib():
    L = ol.oracle()
    L.orc_mac_apply.argtypes = [I, I, D, D, ol._dp, ol._dp, ol._dp, I, I, I]
    L.orc_fem_apply.argtypes = [I, I, D, D, ol._dp, ol._dp, ol._dp, I, I, I]
    L.orc_mac_solve.argtypes = [I, I, D, D, ol._dp, ol._dp, ol._dp, D, ol._dp, I, I, I]
    L.orc_fem_solve.argtypes = [I, I, D, D, ol._dp, ol._dp, ol._dp, D, ol._dp, I, I, I]
    return L


def _ptr(a):
    return None if a is None else ol.dptr(a)


def mac_apply(m, n, dx, dy, irho, x, px=False, py=False, zpr=False):
    y = np.zeros(m * n)
    _lib().orc_mac_apply(m, n, 1 / dx**2, 1 / dy**2, _ptr(irho), ol.dptr(x), ol.dptr(y), int(px), int(py), int(zpr))
    return y


def fem_apply(m, n, dx, dy, irho, x, px=False, py=False, zpr=False):
    y = np.zeros(m * n)
    _lib().orc_fem_apply(m, n, dx, dy, _ptr(irho), ol.dptr(x), ol.dptr(y), int(px), int(py), int(zpr))
    return y


def mac_solve(m, n, dx, dy, irho, b, tol=1e-13, px=False, py=False, zpr=False):
    z = np.zeros(m * n)
    r = D(0)
    c = _lib().orc_mac_solve(m, n, 1 / dx**2, 1 / dy**2, _ptr(irho), ol.dptr(b), ol.dptr(z), tol, ctypes.byref(r),
                             int(px), int(py), int(zpr))
    return z, c, r.value


def fem_solve(m, n, dx, dy, irho, b, tol=1e-13, px=False, py=False, zpr=False):
    z = np.zeros(m * n)
    r = D(0)
    c = _lib().orc_fem_solve(m, n, dx, dy, _ptr(irho), ol.dptr(b), ol.dptr(z), tol, ctypes.byref(r),
                             int(px), int(py), int(zpr))
    return z, c, r.value


SIZES = [(8, 8), (16, 12), (7, 5), (32, 32), (24, 40)]


@pytest.mark.parametrize("m,n", SIZES)
@pytest.mark.parametrize("varying", [False, True])
def test_mac_operator_is_face_flux_laplacian(m, n, varying):
    dx, dy = 2.0 / m, 1.5 / n
    irho = ops.mac_irho_from_rho(ops.blob_density(m, n, seed=1), m, n) if varying else None
    A = ops.mac_matrix(m, n, dx, dy, irho)
    x = np.random.default_rng(2).standard_normal(m * n)
    y = mac_apply(m, n, dx, dy, irho, x)
    assert np.abs(y - A @ x).max() <= 1e-13 * np.abs(A @ x).max()
    assert np.abs(mac_apply(m, n, dx, dy, irho, np.ones(m * n))).max() < 1e-10  # constants are the null space


@pytest.mark.parametrize("m,n", [(9, 9), (17, 13), (6, 5), (33, 33), (25, 41)])
@pytest.mark.parametrize("varying", [False, True])
def test_fem_operator_is_bilinear_stiffness(m, n, varying):
    dx, dy = 1.0 / (m - 1), 0.7 / (n - 1)
    irho = ops.fem_irho_from_rho(ops.blob_density(m - 1, n - 1, seed=3), m - 1, n - 1) if varying else None
    A = ops.fem_matrix(m, n, dx, dy, irho)
    x = np.random.default_rng(4).standard_normal(m * n)
    y = fem_apply(m, n, dx, dy, irho, x)
    assert np.abs(y - A @ x).max() <= 1e-13 * np.abs(A @ x).max()
    assert np.abs(A - A.T).max() < 1e-14  # Galerkin stiffness is symmetric


@pytest.mark.parametrize("m,n", [(16, 16), (12, 20), (32, 32)])
@pytest.mark.parametrize("kind", ["mac", "fem"])
def test_solve_matches_direct(m, n, kind):
    rng = np.random.default_rng(5)
    if kind == "mac":
        dx, dy = 1.0 / m, 1.0 / n
        irho = ops.mac_irho_from_rho(ops.blob_density(m, n, seed=6), m, n)
        A = ops.mac_matrix(m, n, dx, dy, irho)
        b = rng.standard_normal(m * n)
        z, cyc, rr = mac_solve(m, n, dx, dy, irho, b - b.mean())
    else:
        mf, nf = m + 1, n + 1
        dx, dy = 1.0 / m, 1.0 / n
        irho = ops.fem_irho_from_rho(ops.blob_density(m, n, seed=6), m, n)
        A = ops.fem_matrix(mf, nf, dx, dy, irho)
        b = rng.standard_normal(mf * nf)
        z, cyc, rr = fem_solve(mf, nf, dx, dy, irho, b - b.mean())
    assert cyc > 0 and rr <= 1e-13
    zd = ops.neumann_solve(A, b)
    assert np.linalg.norm((z - z.mean()) - zd) <= 1e-10 * np.linalg.norm(zd)


def _mms_error_mac(m):
    # q = cos(pi x) cos(pi y) on [0,1]^2 has zero normal derivative on the walls
    h = 1.0 / m
    x = (np.arange(m) + 0.5) * h
    X, Y = np.meshgrid(x, x)
    q = np.cos(np.pi * X) * np.cos(np.pi * Y)
    b = (-2 * np.pi**2 * q).ravel()
    z, _, _ = mac_solve(m, m, h, h, None, b - b.mean())
    return np.abs((z - z.mean()) - (q.ravel() - q.mean())).max()


def _mms_error_fem(m):
    h = 1.0 / m
    x = np.arange(m + 1) * h
    X, Y = np.meshgrid(x, x)
    q = np.cos(np.pi * X) * np.cos(np.pi * Y)
    # the load of -lap q = 2 pi^2 q against each hat function, by 3x3 Gauss
    # quadrature per cell (exact enough for the order test)
    A = ops.fem_matrix(m + 1, m + 1, h, h)
    load = np.zeros((m + 1) * (m + 1))
    gp, gw = np.polynomial.legendre.leggauss(3)
    for cj in range(m):
        for ci in range(m):
            for a, wa in zip(gp, gw):
                for bq, wb in zip(gp, gw):
                    s, t = 0.5 * (a + 1), 0.5 * (bq + 1)
                    xx, yy = (ci + s) * h, (cj + t) * h
                    f = 2 * np.pi**2 * np.cos(np.pi * xx) * np.cos(np.pi * yy) * 0.25 * wa * wb * h * h
                    for (di, dj, phi) in ((0, 0, (1 - s) * (1 - t)), (1, 0, s * (1 - t)), (0, 1, (1 - s) * t), (1, 1, s * t)):
                        load[(cj + dj) * (m + 1) + ci + di] += f * phi
    z, _, _ = fem_solve(m + 1, m + 1, h, h, None, load - load.mean())
    qq = q.ravel()
    del A
    return np.abs((z - z.mean()) - (qq - qq.mean())).max()


@pytest.mark.parametrize("kind", ["mac", "fem"])
def test_manufactured_solution_second_order(kind):
    err = [(_mms_error_mac if kind == "mac" else _mms_error_fem)(m) for m in (16, 32, 64)]
    rates = [np.log2(err[k] / err[k + 1]) for k in range(2)]
    assert all(1.8 < r < 2.3 for r in rates), (err, rates)


PERIODS = [(True, False), (False, True), (True, True)]


@pytest.mark.parametrize("px,py", PERIODS)
@pytest.mark.parametrize("m,n", [(16, 12), (24, 8)])
@pytest.mark.parametrize("varying", [False, True])
def test_periodic_operators_match_definitions(px, py, m, n, varying):
    """Periodic MAC faces wrap across the seam (its right side the
    right-edge irho entries); periodic FEM directions have m nodes and m
    cells, the seam cell joining node m-1 to node 0."""
    rng = np.random.default_rng(7)
    dx, dy = 2.0 / m, 1.5 / n
    R = ops.blob_density(m, n, seed=3)  # not periodic: the seam faces differ
    irho = ops.mac_irho_from_rho(R, m, n) if varying else None
    A = ops.mac_matrix(m, n, dx, dy, irho, px, py)
    x = rng.standard_normal(m * n)
    assert np.abs(mac_apply(m, n, dx, dy, irho, x, px, py) - A @ x).max() <= 1e-12 * np.abs(A @ x).max()
    mf, nf = (m if px else m + 1), (n if py else n + 1)
    hx, hy = 2.0 / m, 1.5 / n
    irf = ops.fem_irho_periodic(R, mf, nf, px, py) if varying else None
    F = ops.fem_matrix(mf, nf, hx, hy, irf, px, py)
    x = rng.standard_normal(mf * nf)
    assert np.abs(fem_apply(mf, nf, hx, hy, irf, x, px, py) - F @ x).max() <= 1e-12 * np.abs(F @ x).max()


@pytest.mark.parametrize("px,py", PERIODS)
@pytest.mark.parametrize("kind", ["mac", "fem"])
def test_periodic_solves_match_direct(px, py, kind):
    m, n = 32, 16
    R = ops.periodic_density(m, n, px, py, seed=5)
    rng = np.random.default_rng(8)
    if kind == "mac":
        dx, dy = 2.0 / m, 1.5 / n
        irho = ops.mac_irho_from_rho(R, m, n)
        A = ops.mac_matrix(m, n, dx, dy, irho, px, py)
        b = rng.standard_normal(m * n)
        z, cyc, rr = mac_solve(m, n, dx, dy, irho, b - b.mean(), 1e-12, px, py)
    else:
        mf, nf = (m if px else m + 1), (n if py else n + 1)
        dx, dy = 2.0 / m, 1.5 / n
        irho = ops.fem_irho_periodic(R, mf, nf, px, py)
        A = ops.fem_matrix(mf, nf, dx, dy, irho, px, py)
        b = rng.standard_normal(mf * nf)
        z, cyc, rr = fem_solve(mf, nf, dx, dy, irho, b - b.mean(), 1e-12, px, py)
    assert cyc > 0 and rr <= 1e-12
    zd = ops.neumann_solve(A, b)
    assert np.abs((z - z.mean()) - zd).max() <= 1e-9 * np.abs(zd).max()


def test_periodic_poisson_second_order():
    """Doubly periodic MAC Poisson problem with the exact solution
    sin(pi x) cos(2 pi y) on [-1,1]^2: the error falls by about 4 per
    halving of h."""
    errs = []
    for m in (16, 32, 64):
        h = 2.0 / m
        x = -1 + (np.arange(m) + 0.5) * h
        X, Y = np.meshgrid(x, x)
        ue = np.sin(np.pi * X) * np.cos(2 * np.pi * Y)
        f = -5 * np.pi**2 * ue
        z, _, _ = mac_solve(m, m, h, h, None, f.ravel() - f.mean(), 1e-13, True, True)
        z = z.reshape(m, m)
        errs.append(np.abs((z - z.mean()) - (ue - ue.mean())).max())
    r = [errs[k] / errs[k + 1] for k in range(2)]
    assert all(3.6 < q < 4.4 for q in r), r


@pytest.mark.parametrize("kind", ["mac", "fem"])
@pytest.mark.parametrize("m", [10, 15, 35])
def test_periodic_resampled_solves_match_direct(kind, m):
    """Doubly periodic sizes whose halving does not end on a tiny level
    (10 -> 5, 35 odd, 15 odd: conv_test's ladder halves to 35, 45, 63, 105
    or 315) coarsen through a resampled 2^k / 3 2^k level; odd sizes are
    swept lexicographically. The solve still equals the direct one."""
    n = m
    R = ops.periodic_density(m, n, True, True, seed=6)
    rng = np.random.default_rng(9)
    dx, dy = 2.0 / m, 2.0 / n
    if kind == "mac":
        irho = ops.mac_irho_from_rho(R, m, n)
        A = ops.mac_matrix(m, n, dx, dy, irho, True, True)
        b = rng.standard_normal(m * n)
        z, cyc, rr = mac_solve(m, n, dx, dy, irho, b - b.mean(), 1e-12, True, True)
    else:
        irho = ops.fem_irho_periodic(R, m, n, True, True)
        A = ops.fem_matrix(m, n, dx, dy, irho, True, True)
        b = rng.standard_normal(m * n)
        z, cyc, rr = fem_solve(m, n, dx, dy, irho, b - b.mean(), 1e-12, True, True)
    assert 0 < cyc < 60 and rr <= 1e-12, (cyc, rr)
    zd = ops.neumann_solve(A, b)
    assert np.abs((z - z.mean()) - zd).max() <= 1e-9 * np.abs(zd).max()
