"""Periodic projection solvers (the x_prd / y_prd forms of mgs_mac_* and
mgs_fem_*, mgs_mac.hh:45-56, mgs_mac.cc:52-53, mgs_fem.hh:47-52,
mgs_fem.cc:60-76, 140-160). Pinned to the operators assembled in
tests/operators.py from their definitions (MAC faces wrapping across the
seam, FEM bilinear elements with the seam cell joining node m-1 to node 0):
the GPU operator equals them to 1e-13, a solve to relative residual 1e-10
equals the mean-free direct solution to 1e-8, every fused smoother launch
shape gives the per-colour sweep bit for bit when the seam faces agree,
and unsupported sizes are refused."""
import numpy as np
import pytest

import incrmt_amd as ia
from tests import operators as ops

pytestmark = pytest.mark.gpu

PERIODS = [(True, False), (False, True), (True, True)]


def _mac(m, n, px, py, varying, seed=1):
    dx, dy = 2.0 / m, 1.5 / n
    irho = None
    if varying:
        irho = ops.mac_irho_from_rho(ops.periodic_density(m, n, px, py, seed=seed), m, n)
    s = ia.MgSolver(ia.MG_MAC, m, n, dx, dy, x_prd=px, y_prd=py)
    s.set_irho(irho)
    return s, ops.mac_matrix(m, n, dx, dy, irho, px, py)


def _fem(m, n, px, py, varying, seed=1):
    dx = 2.0 / (m if px else m - 1)
    dy = 1.5 / (n if py else n - 1)
    irho = None
    if varying:
        R = ops.periodic_density(m, n, px, py, seed=seed)
        irho = ops.fem_irho_periodic(R, m, n, px, py)
    s = ia.MgSolver(ia.MG_FEM, m, n, dx, dy, x_prd=px, y_prd=py)
    s.set_irho(irho)
    return s, ops.fem_matrix(m, n, dx, dy, irho, px, py)


def _sizes(kind, px, py, m, n):
    """Node counts along non-periodic FEM directions are 2^k+1."""
    if kind == "fem":
        return (m if px else m + 1), (n if py else n + 1)
    return m, n


@pytest.mark.parametrize("kind", ["mac", "fem"])
@pytest.mark.parametrize("px,py", PERIODS)
@pytest.mark.parametrize("varying", [False, True])
def test_periodic_operator(kind, px, py, varying):
    m, n = _sizes(kind, px, py, 32, 16)
    s, A = (_mac if kind == "mac" else _fem)(m, n, px, py, varying)
    x = np.random.default_rng(0).standard_normal(m * n)
    y = s.apply(x)
    assert np.abs(y - A @ x).max() <= 1e-13 * np.abs(A @ x).max()
    s.close()


@pytest.mark.parametrize("kind", ["mac", "fem"])
@pytest.mark.parametrize("px,py", PERIODS)
@pytest.mark.parametrize("mn", [(64, 32), (48, 48)])
@pytest.mark.parametrize("precision", [ia.PREC_FP64, ia.PREC_MIXED])
def test_periodic_solve_matches_direct(kind, px, py, mn, precision):
    m, n = _sizes(kind, px, py, *mn)
    s, A = (_mac if kind == "mac" else _fem)(m, n, px, py, True, seed=2)
    b = np.random.default_rng(3).standard_normal(m * n)
    z, cyc, rr = s.solve(b, rel_tol=1e-10, max_cycles=200, precision=precision)
    print(kind, px, py, m, n, precision, cyc, rr, flush=True)
    assert rr <= 1e-10
    zd = ops.neumann_solve(A, b)
    zz = z - z.mean()
    assert np.abs(zz - zd).max() <= 1e-8 * np.abs(zd).max()
    s.close()


@pytest.mark.parametrize("kind,shapes", [("mac", [0, 1, 2, 3, 4, 5, 8, 9]), ("fem", [0, 1, 2, 6, 7, 8, 9, 10, 11, 12])])
@pytest.mark.parametrize("px,py", PERIODS)
def test_periodic_shapes_bit_identical(kind, shapes, px, py):
    m, n = _sizes(kind, px, py, 128, 64)
    s, _ = (_mac if kind == "mac" else _fem)(m, n, px, py, True, seed=4)
    rng = np.random.default_rng(5)
    b, z0 = rng.standard_normal(m * n), rng.standard_normal(m * n)
    ref = s.smooth(b, z0, 4, -1)
    for sh in shapes:
        got = s.smooth(b, z0, 4, sh)
        assert np.array_equal(got, ref), sh
    s.close()


def test_periodic_graph_equals_direct():
    for kind in ("mac", "fem"):
        g, _ = (_mac if kind == "mac" else _fem)(128, 64, True, True, True)
        d, _ = (_mac if kind == "mac" else _fem)(128, 64, True, True, True)
        g.set_launch(ia.LAUNCH_GRAPH)
        d.set_launch(ia.LAUNCH_DIRECT)
        rng = np.random.default_rng(6)
        for rep in range(3):
            b = rng.standard_normal(128 * 64)
            zg, cg, _ = g.solve(b, rel_tol=1e-10, precision=ia.PREC_MIXED)
            zd, cd, _ = d.solve(b, rel_tol=1e-10, precision=ia.PREC_MIXED)
            assert cg == cd and np.array_equal(zg, zd), (kind, rep)
        g.close()
        d.close()


@pytest.mark.parametrize("kind,m,n", [("mac", 30, 16), ("mac", 80, 80), ("fem", 31, 16), ("fem", 40, 40)])
def test_periodic_unsupported_sizes(kind, m, n):
    """One periodic direction: odd sizes along it, and sizes whose
    halvings end above 16 points."""
    with pytest.raises(ia.IncrmtError) as e:
        ia.MgSolver(ia.MG_MAC if kind == "mac" else ia.MG_FEM, m, n, 0.1, 0.1, x_prd=True, y_prd=False)
    assert e.value.code == ia.ERR_UNSUPPORTED
    with pytest.raises(ia.IncrmtError):  # too small to resample
        ia.MgSolver(ia.MG_MAC if kind == "mac" else ia.MG_FEM, 7, 7, 0.1, 0.1, x_prd=True, y_prd=True)


@pytest.mark.parametrize("kind", ["mac", "fem"])
@pytest.mark.parametrize("m", [10, 15, 35, 45, 63])
@pytest.mark.parametrize("precision", [ia.PREC_FP64, ia.PREC_MIXED])
def test_doubly_periodic_resampled_matches_direct(kind, m, precision):
    """Doubly periodic sizes whose halving does not end on a tiny level
    (10 -> 5; 35, 45, 63: where conv_test's 280, 360 and 504 halve to; 15
    odd, smoothed in nine colours) coarsen through a resampled level: the
    solve still equals the mean-free direct solution."""
    s, A = (_mac if kind == "mac" else _fem)(m, m, True, True, True, seed=4)
    b = np.random.default_rng(5).standard_normal(m * m)
    z, cyc, rr = s.solve(b, rel_tol=1e-10, max_cycles=200, precision=precision)
    print(kind, m, precision, cyc, rr, flush=True)
    assert rr <= 1e-10 and cyc < 40
    zd = ops.neumann_solve(A, b)
    assert np.abs((z - z.mean()) - zd).max() <= 1e-8 * np.abs(zd).max()
    s.close()


@pytest.mark.parametrize("px,py", PERIODS)
def test_periodic_seam_faces_differ(px, py):
    """A density that is not periodic across the seam (a solid crossing it)
    gives the seam two different MAC faces: the left-edge entry left of
    cell 0 and the right-edge entry right of cell m-1 (mgs_mac.hh:81-96).
    Every fused launch shape still gives the per-colour sweep, which uses
    the operator's own faces, bit for bit."""
    m, n = 128, 64
    dx, dy = 2.0 / m, 1.5 / n
    irho = ops.mac_irho_from_rho(ops.blob_density(m, n, 3.0, seed=7), m, n)
    s = ia.MgSolver(ia.MG_MAC, m, n, dx, dy, x_prd=px, y_prd=py)
    s.set_irho(irho)
    rng = np.random.default_rng(8)
    b, z0 = rng.standard_normal(m * n), rng.standard_normal(m * n)
    ref = s.smooth(b, z0, 4, -1)
    for sh in [0, 1, 2, 3, 4, 5, 8, 9]:
        got = s.smooth(b, z0, 4, sh)
        assert np.array_equal(got, ref), sh
    s.close()
