"""GPU projection solvers (mgs_mac_* / mgs_fem_* replacements).

Operators: equal to the definitions assembled in tests/operators.py to
1e-13 relative (the same operators the oracle is pinned to). Solves: the
relative residual reaches the requested tolerance, and the mean-free
solution agrees with a direct solve / the oracle to 1e-8 relative at
tolerance 1e-10. The reference's own solver (TGMG) is not in the reference
tree: its smoother, coarse operators and cycle counts are not reproduced,
only the solution it is asked to produce.
"""
import math

import numpy as np
import pytest

import incrmt_amd as ia
from tests import operators as ops

pytestmark = pytest.mark.gpu


def _mac(m, n, varying, seed=1):
    dx, dy = 1.0 / m, 1.3 / n
    irho = ops.mac_irho_from_rho(ops.blob_density(m, n, seed=seed), m, n) if varying else None
    s = ia.MgSolver(ia.MG_MAC, m, n, dx, dy)
    s.set_irho(irho)
    return s, ops.mac_matrix(m, n, dx, dy, irho)


def _fem(m, n, varying, seed=1):
    dx, dy = 1.0 / (m - 1), 1.3 / (n - 1)
    irho = ops.fem_irho_from_rho(ops.blob_density(m - 1, n - 1, seed=seed), m - 1, n - 1) if varying else None
    s = ia.MgSolver(ia.MG_FEM, m, n, dx, dy)
    s.set_irho(irho)
    return s, ops.fem_matrix(m, n, dx, dy, irho)


@pytest.mark.parametrize("m,n", [(16, 16), (64, 40), (30, 18), (7, 9)])
@pytest.mark.parametrize("varying", [False, True])
def test_mac_operator(m, n, varying):
    s, A = _mac(m, n, varying)
    x = np.random.default_rng(0).standard_normal(m * n)
    assert np.abs(s.apply(x) - A @ x).max() <= 1e-13 * np.abs(A @ x).max()


@pytest.mark.parametrize("m,n", [(17, 17), (65, 41), (31, 19), (8, 9)])
@pytest.mark.parametrize("varying", [False, True])
def test_fem_operator(m, n, varying):
    s, A = _fem(m, n, varying)
    x = np.random.default_rng(0).standard_normal(m * n)
    assert np.abs(s.apply(x) - A @ x).max() <= 1e-13 * np.abs(A @ x).max()


@pytest.mark.parametrize("kind", ["mac", "fem"])
@pytest.mark.parametrize("m,n", [(32, 32), (48, 20), (33, 17)])
@pytest.mark.parametrize("precision", [ia.PREC_FP64, ia.PREC_MIXED])
def test_solve_matches_direct(kind, m, n, precision):
    s, A = (_mac if kind == "mac" else _fem)(m, n, True, seed=2)
    b = np.random.default_rng(3).standard_normal(m * n)
    z, cyc, rr = s.solve(b, rel_tol=1e-10, max_cycles=400, precision=precision)
    assert rr <= 1e-10
    zd = ops.neumann_solve(A, b)
    assert np.linalg.norm((z - z.mean()) - zd) <= 1e-8 * np.linalg.norm(zd)


@pytest.mark.parametrize("kind", ["mac", "fem"])
def test_warm_start_and_zero_rhs(kind):
    s, _ = (_mac if kind == "mac" else _fem)(64, 64, True)
    b = np.random.default_rng(4).standard_normal(64 * 64)
    z1, c1, _ = s.solve(b, rel_tol=1e-10)
    z2, c2, r2 = s.solve(b, rel_tol=1e-10)  # warm start from z1
    assert c1 > 0 and c2 == 0 and r2 <= 1e-10
    z3, c3, r3 = s.solve(np.zeros(64 * 64), rel_tol=1e-10)
    assert c3 == 0


def test_not_converged_is_reported():
    s, _ = _mac(64, 64, True)
    b = np.random.default_rng(5).standard_normal(64 * 64)
    with pytest.raises(ia.IncrmtError) as e:
        s.solve(b, rel_tol=1e-15, max_cycles=1)
    assert e.value.code == ia.ERR_NOT_CONVERGED


def fourier_rhs(nodes, seed=20260101):
    """The standalone FEM configuration's right-hand side (incrmt_amd.workloads):
    the FEM source (fluid_2d.cc:1303-1308, dt = 1) of a seeded smooth velocity
    of 32 Fourier modes with amplitudes ~ |k|^-2; mean projected out."""
    from incrmt_amd.workloads import fourier_fem_rhs
    b, h = fourier_fem_rhs(nodes, seed)
    b = b.ravel()
    return b - b.mean()


def _fem_residual(z, b, n, h):
    # independent check of ||b - A z|| / ||b|| with a sparse assembly of the
    # bilinear-FEM operator (the definition tests/operators.py assembles)
    A = ops.fem_matrix_sparse(n, n, h, h)
    return np.linalg.norm(b - A @ z) / np.linalg.norm(b)


@pytest.mark.parametrize("precision", [ia.PREC_FP64, ia.PREC_MIXED])
def test_fem_1024_fourier_projection(precision):
    # the standalone FEM projection on 1024x1024 cells (1025^2 nodes, the
    # step's FEM size): converges to a relative residual of 1e-10 from zero
    n = 1025
    s = ia.MgSolver(ia.MG_FEM, n, n, 1.0 / (n - 1), 1.0 / (n - 1))
    b = fourier_rhs(n)
    z, cyc, rr = s.solve(b, rel_tol=1e-10, max_cycles=100, precision=precision)
    assert rr <= 1e-10 and cyc < 40
    assert np.isfinite(z).all()


@pytest.mark.parametrize("m,n", [(34, 34), (40, 28), (64, 33), (20, 64)])
@pytest.mark.parametrize("varying", [False, True])
def test_fem_even_nodes_matches_direct(m, n, varying):
    # node counts vertex coarsening cannot halve: the first coarse level is a
    # uniform resampling of the box (non-nested transfers); the solution is
    # still the fine system's to the requested tolerance
    s, A = _fem(m, n, varying, seed=3)
    b = np.random.default_rng(9).standard_normal(m * n)
    b -= b.mean()
    for precision in (ia.PREC_FP64, ia.PREC_MIXED):
        s.z[:] = 0
        z, cyc, rr = s.solve(b, rel_tol=1e-10, max_cycles=200, precision=precision)
        assert rr <= 1e-10 and cyc < 60, (precision, cyc, rr)
        zd = ops.neumann_solve(A, b)
        assert np.linalg.norm((z - z.mean()) - zd) <= 1e-8 * np.linalg.norm(zd)


@pytest.mark.parametrize("precision", [ia.PREC_FP64, ia.PREC_MIXED])
def test_fem_1024_nodes_config(precision):
    # the standalone FEM configuration: 1024 x 1024 nodes on the unit square
    # (h = 1/1023), seeded Fourier right-hand side, zero start, relative
    # residual 1e-10, FP64 and FP32-smoothing refinement
    n = 1024
    h = 1.0 / (n - 1)
    s = ia.MgSolver(ia.MG_FEM, n, n, h, h)
    b = fourier_rhs(n)
    z, cyc, rr = s.solve(b, rel_tol=1e-10, max_cycles=100, precision=precision)
    print(f"1024^2 nodes precision={precision}: {cyc} V-cycles, rel. residual {rr:.3e}", flush=True)
    assert rr <= 1e-10 and cyc < 40
    assert _fem_residual(z, b, n, h) <= 2e-10


@pytest.mark.parametrize("kind", ["mac", "fem"])
@pytest.mark.parametrize("sweeps", [1, 2, 3])
@pytest.mark.parametrize("merged", [False, True])
def test_fused_smoother_shapes_bit_identical(kind, sweeps, merged):
    # every fused launch shape the autotuner may pick gives exactly the
    # sweep of one launch per colour (the reference form of the sweep);
    # merged: sizes one point past a multiple of the tile widths, where the
    # last row and column of tiles fold into the halo of the ones before
    if merged:
        m, n = (129, 65) if kind == "mac" else (257, 129)
    else:
        m, n = (200, 136) if kind == "mac" else (201, 137)
    s, _ = (_mac if kind == "mac" else _fem)(m, n, True, seed=5)
    rng = np.random.default_rng(6)
    b = rng.standard_normal(m * n)
    z = rng.standard_normal(m * n)
    ref = s.smooth(b, z, sweeps, -1)
    shapes = (0, 1, 2, 3, 4, 5, 8, 9, 14) if kind == "mac" else (0, 1, 2, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)
    for sh in shapes:
        assert np.array_equal(s.smooth(b, z, sweeps, sh), ref), sh
    with pytest.raises(ia.IncrmtError):
        s.smooth(b, z, sweeps, 6 if kind == "mac" else 3)
    if kind == "fem":
        with pytest.raises(ia.IncrmtError):  # fits in LDS only in single precision
            s.smooth(b, z, sweeps, 16)
    # single precision: every shape, including those only FP32 tiles fit,
    # gives the per-colour FP32 sweep bit for bit
    ref32 = s.smooth(b, z, sweeps, -1, precision=ia.PREC_MIXED)
    assert np.abs(ref32 - ref).max() <= 1e-4 * np.abs(ref).max()
    shapes32 = shapes + (() if kind == "mac" else (16,))
    for sh in shapes32:
        assert np.array_equal(s.smooth(b, z, sweeps, sh, precision=ia.PREC_MIXED), ref32), ("fp32", sh)


@pytest.mark.parametrize("precision", [ia.PREC_FP64, ia.PREC_MIXED])
def test_time_solve_on_device(precision):
    """The device-timed solve (b kept on the device, zero start) runs the
    same solve as incrmt_mg_solve from a zero start: same V-cycles, same
    residual, a positive event time."""
    m = 129
    s, _ = _fem(m, m, True, seed=3)
    b = np.random.default_rng(4).standard_normal(m * m)
    z, cyc, rr = s.solve(b, rel_tol=1e-10, max_cycles=100, precision=precision, z=np.zeros(m * m))
    ms, c2, r2 = s.time_solve(1e-10, 100, precision, reps=2)
    assert c2 == cyc and r2 == rr and ms > 0
    s.close()
