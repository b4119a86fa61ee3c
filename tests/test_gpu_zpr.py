"""Horizontal flow with zero pressure on the right edge (sim_horiz_flow,
fluid_2d::zpr) on the GPU.

Solvers: the MAC system with the right edge face held at q = 0
(mgs_mac.hh:50, 95) and the FEM system with node column m-1 fixed at zero
(mgs_fem.cc:21, 85, 112), pinned to the matrices assembled from their
definitions in tests/operators.py (which the oracle tests in turn pin to
closed-form eigenvectors): the GPU operator equals them to 1e-13, a solve
to relative residual 1e-10 equals the sparse direct solution to 1e-8
(the systems are definite: no mean is removed), every fused smoother
launch shape gives the per-colour sweep bit for bit, and graph replays
equal direct launches.

Step: fluid_2d::step_forward with the inflow ghost columns
(sim_type.cc:51-55), the extrapolated outflow columns (fluid_2d.cc:701-704),
the edge conditions of fluid_2d.cc:745-749 and 847-851 and no pressure
average (fluid_2d.cc:1318), against the oracle per step with the no-slip
tolerances (u, v, p relative L2 <= 1e-6; X, Y within 1e-9; band status
equal on 99.5% of the band, phi within 0.05 dx)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

import incrmt_amd as ia
from tests import operators as ops
from tests import oracle_lib as ol
from tests.test_gpu_step import compare

pytestmark = pytest.mark.gpu


def _mac(m, n, py, varying, seed=1):
    dx, dy = 4.0 / m, 2.0 / n
    # a y seam needs the same density on both of its sides (the fused
    # smoother uses one face for both, see tests/test_gpu_mg_periodic.py)
    irho = ops.mac_irho_from_rho(ops.periodic_density(m, n, False, py, seed=seed), m, n) if varying else None
    s = ia.MgSolver(ia.MG_MAC, m, n, dx, dy, y_prd=py, zpr=True)
    s.set_irho(irho)
    return s, ops.mac_matrix(m, n, dx, dy, irho, py=py, zpr=True)


def _fem(m, n, py, varying, seed=1):
    """m nodes per row, the last one fixed; n nodes per column."""
    dx, dy = 4.0 / (m - 1), 2.0 / (n if py else n - 1)
    irho = None
    if varying:
        R = ops.periodic_density(m, n, False, py, seed=seed)
        irho = np.zeros(m * n + m)
        irho[: m * n] = (1.0 / R[0:n, 0:m]).ravel()
    s = ia.MgSolver(ia.MG_FEM, m, n, dx, dy, y_prd=py, zpr=True)
    s.set_irho(irho)
    return s, ops.fem_matrix(m, n, dx, dy, None if irho is None else irho[: m * n], py=py, zpr=True)


def _sizes(kind, py, m, n):
    if kind == "fem":
        return m + 1, (n if py else n + 1)
    return m, n


def _fixed_zero(kind, m, n, x):
    if kind == "fem":
        x = x.reshape(n, m).copy()
        x[:, -1] = 0.0
        return x.ravel()
    return x


@pytest.mark.parametrize("kind", ["mac", "fem"])
@pytest.mark.parametrize("py", [True, False])
@pytest.mark.parametrize("varying", [False, True])
def test_zpr_operator(kind, py, varying):
    m, n = _sizes(kind, py, 64, 32)
    s, A = (_mac if kind == "mac" else _fem)(m, n, py, varying)
    x = _fixed_zero(kind, m, n, np.random.default_rng(0).standard_normal(m * n))
    y = s.apply(x)
    assert np.abs(y - A @ x).max() <= 1e-13 * np.abs(A @ x).max()
    s.close()


@pytest.mark.parametrize("kind", ["mac", "fem"])
@pytest.mark.parametrize("py", [True, False])
@pytest.mark.parametrize("mn", [(64, 32), (48, 48)])
@pytest.mark.parametrize("precision", [ia.PREC_FP64, ia.PREC_MIXED])
def test_zpr_solve_matches_direct(kind, py, mn, precision):
    m, n = _sizes(kind, py, *mn)
    s, A = (_mac if kind == "mac" else _fem)(m, n, py, True, seed=2)
    b = _fixed_zero(kind, m, n, np.random.default_rng(3).standard_normal(m * n))
    z, cyc, rr = s.solve(b, rel_tol=1e-10, max_cycles=200, precision=precision)
    print(kind, py, m, n, precision, cyc, rr, flush=True)
    assert rr <= 1e-10
    zd = spl.spsolve(A.tocsc(), b)
    assert np.abs(z - zd).max() <= 1e-8 * np.abs(zd).max()
    if kind == "fem":
        assert np.all(z.reshape(n, m)[:, -1] == 0.0)
    s.close()


@pytest.mark.parametrize("kind,shapes", [("mac", [0, 1, 2, 3, 4, 5, 8, 9]), ("fem", [0, 1, 2, 6, 7, 8, 9, 10, 11, 12])])
@pytest.mark.parametrize("py", [True, False])
def test_zpr_shapes_bit_identical(kind, shapes, py):
    m, n = _sizes(kind, py, 128, 64)
    s, _ = (_mac if kind == "mac" else _fem)(m, n, py, True, seed=4)
    rng = np.random.default_rng(5)
    b = _fixed_zero(kind, m, n, rng.standard_normal(m * n))
    z0 = _fixed_zero(kind, m, n, rng.standard_normal(m * n))
    ref = s.smooth(b, z0, 4, -1)
    for sh in shapes:
        got = s.smooth(b, z0, 4, sh)
        assert np.array_equal(got, ref), sh
    s.close()


def test_zpr_graph_equals_direct():
    for kind in ("mac", "fem"):
        m, n = _sizes(kind, True, 128, 64)
        g, _ = (_mac if kind == "mac" else _fem)(m, n, True, True)
        d, _ = (_mac if kind == "mac" else _fem)(m, n, True, True)
        g.set_launch(ia.LAUNCH_GRAPH)
        d.set_launch(ia.LAUNCH_DIRECT)
        rng = np.random.default_rng(6)
        for rep in range(3):
            b = _fixed_zero(kind, m, n, rng.standard_normal(m * n))
            zg, cg, _ = g.solve(b, rel_tol=1e-10, precision=ia.PREC_MIXED)
            zd, cd, _ = d.solve(b, rel_tol=1e-10, precision=ia.PREC_MIXED)
            assert cg == cd and np.array_equal(zg, zd), (kind, rep)
        g.close()
        d.close()


def test_zpr_bad_arguments():
    with pytest.raises(ia.IncrmtError) as e:
        ia.MgSolver(ia.MG_MAC, 64, 32, 0.1, 0.1, x_prd=True, zpr=True)
    assert e.value.code == ia.ERR_ARG
    with pytest.raises(ia.IncrmtError) as e:
        ia.MgSolver(ia.MG_MAC, 80, 80, 0.1, 0.1, y_prd=True, zpr=True)
    assert e.value.code == ia.ERR_UNSUPPORTED
    f = ia.Fluid2D(64, 64, visc=1e-3, rhof=1.0, x_prd=True, y_prd=True)
    with pytest.raises(ia.IncrmtError) as e:
        f.set_horizontal_flow(1.0)
    assert e.value.code == ia.ERR_ARG
    f.close()


def gpu_from_spec(spec, tol=1e-12):
    f = ia.Fluid2D(spec["m"], spec["n"], visc=spec["visc"], rhof=spec["rhof"], tw=spec["tw"], pinw=spec["pinw"],
                   ax=spec["ax"], bx=spec["bx"], ay=spec["ay"], by=spec["by"], x_prd=False,
                   y_prd=bool(spec["y_prd"]), w_scale=spec["w_scale"], w_force=spec["w_force"])
    for kind, p, G, rhos, evm in spec["objects"]:
        f.add_circle(p[0], p[1], p[2], p[3], G=G, rhos=rhos, ev_trans_mult=evm)
    f.set_horizontal_flow(spec["U"])
    f.set_solver(rel_tol=tol, max_cycles=200)
    f.initialize(spec["ev_mult"], spec["dt_pad"], spec["dt_ev_pad"])
    return f


@pytest.mark.parametrize("m,n,y_prd,steps", [(128, 64, True, 10), (64, 32, False, 8)])
def test_horiz_flow_per_step(m, n, y_prd, steps):
    spec = ol.horiz_flow_spec(m, n, U=1.0, y_prd=y_prd)
    g, o = gpu_from_spec(spec), ol.OracleFluid(spec, mg_tol=1e-13)
    assert g.dt_reg == pytest.approx(o.dt_reg, rel=1e-15)
    dx = 4.0 / m
    compare(g, o, m, dx, "init")
    for k in range(1, steps + 1):
        g.step_forward(0.2 * dx)
        o.step(0.2 * dx)
        compare(g, o, m, dx, f"step {k}")
    u, v, p = g.get("u"), g.get("v"), g.get("p")
    assert np.all(u[:, 0:2] == 1.0) and np.all(v[:, 0:2] == 0.0)
    assert np.all(p[2:n + 3, 2 + m] == 0.0)
    g.close()
    o.close()


def test_horiz_flow_properties():
    """256x128, 6 steps: the corrected faces of every cell but the last
    column are discretely divergence-free to the solver tolerance (the
    reference corrects no edge face, fluid_2d.cc:1193-1235), the pressure
    node column m is exactly zero and the inflow columns hold (U, 0)."""
    spec = ol.horiz_flow_spec(256, 128, U=1.0)
    g = gpu_from_spec(spec, tol=1e-10)
    h = 4.0 / 256
    for k in range(6):
        g.step_forward(0.2 * h)
        ul, vd = g.get("ul"), g.get("vd")
        div = (ul[2:-2, 3:-1] - ul[2:-2, 2:-2]) / h + (vd[3:-1, 2:-2] - vd[2:-2, 2:-2]) / h
        assert np.abs(div[:, :-1]).max() <= 1e-7 / h, k
        p = g.get("p")
        assert np.all(p[2:-1, 2 + 256] == 0.0)
        assert np.all(g.get("u")[:, 0:2] == 1.0)
    g.close()


def test_uniform_flow_stays_uniform():
    """No solid, y periodic: the uniform inflow is a steady state of the
    horizontal-flow step, kept to the last bit (the oracle's own check in
    tests/test_oracle_zpr.py)."""
    f = ia.Fluid2D(64, 32, visc=1e-2, rhof=1.0, ax=-1.0, bx=3.0, ay=-1.0, by=1.0, x_prd=False, y_prd=True)
    f.set_horizontal_flow(1.0)
    f.set_solver(rel_tol=1e-12, max_cycles=200)
    f.initialize(0.8, 0.4, 0.8)
    for k in range(4):
        f.step_forward(0.01)
    u, v, p = f.get("u"), f.get("v"), f.get("p")
    assert np.all(u[2:-2, 2:-2] == 1.0) and np.all(v[2:-2, 2:-2] == 0.0) and np.all(p[2:-2, 2:-2] == 0.0)
    f.close()
