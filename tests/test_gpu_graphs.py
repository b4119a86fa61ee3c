"""V-cycles replayed from captured HIP graphs against the same V-cycles
launched kernel by kernel. Stated bar: bit-identical iterates, cycle counts
and residuals on every solve — several solves in a row (each one replays
the graph a few times, with the solver's own eager work — residual checks,
host copies — between the replays), solves after smoothing calls that leave
the other iterate buffer current (the graph is captured again), and whole
steps of the fluid solver (level-set and advection kernels between the
projections), in both arithmetics."""
import numpy as np
import pytest

import incrmt_amd as ia

pytestmark = pytest.mark.gpu


def _solver(kind, m, n, mode):
    if kind == "mac":
        s = ia.MgSolver(ia.MG_MAC, m, n, 1.0 / m, 1.3 / n)
    else:
        s = ia.MgSolver(ia.MG_FEM, m, n, 1.0 / (m - 1), 1.3 / (n - 1))
    s.set_launch(mode)
    return s


@pytest.mark.parametrize("kind,m,n", [("mac", 256, 200), ("fem", 257, 201), ("fem", 256, 256)])
@pytest.mark.parametrize("precision", [ia.PREC_FP64, ia.PREC_MIXED])
def test_graph_solves_equal_direct(kind, m, n, precision):
    g, d = _solver(kind, m, n, ia.LAUNCH_GRAPH), _solver(kind, m, n, ia.LAUNCH_DIRECT)
    rng = np.random.default_rng(11)
    for rep in range(4):
        b = rng.standard_normal(m * n)
        # warm starts from the previous solution, which differs per call
        zg, cg, rg = g.solve(b, rel_tol=1e-10, max_cycles=200, precision=precision)
        zd, cd, rd = d.solve(b, rel_tol=1e-10, max_cycles=200, precision=precision)
        print(kind, m, n, precision, rep, cg, rg, flush=True)
        assert cg == cd and cg > 1, (rep, cg, cd)
        assert rg == rd, (rep, rg, rd)
        assert np.array_equal(zg, zd), rep
    g.close()
    d.close()


@pytest.mark.parametrize("kind,m,n,shape", [("mac", 128, 96, 3), ("fem", 129, 97, 1)])
def test_graph_recaptured_after_buffer_swap(kind, m, n, shape):
    """A smoothing call with an odd number of fused launches (shape 3 for MAC
    and shape 1 for FEM take one sweep per launch) leaves the fine level's
    second buffer current; the next solve must not replay the graph made for
    the first buffer."""
    g, d = _solver(kind, m, n, ia.LAUNCH_GRAPH), _solver(kind, m, n, ia.LAUNCH_DIRECT)
    rng = np.random.default_rng(12)
    b = rng.standard_normal(m * n)
    for rep in range(3):
        zg, cg, _ = g.solve(b, rel_tol=1e-9, max_cycles=200)
        zd, cd, _ = d.solve(b, rel_tol=1e-9, max_cycles=200)
        assert cg == cd and np.array_equal(zg, zd), rep
        z0 = rng.standard_normal(m * n)
        sg = g.smooth(b, z0, 1, shape)
        sd = d.smooth(b, z0, 1, shape)
        assert np.array_equal(sg, sd), rep
        b = rng.standard_normal(m * n)
    g.close()
    d.close()


@pytest.mark.parametrize("precision", [ia.PREC_FP64, ia.PREC_MIXED])
@pytest.mark.parametrize("guess", [ia.GUESS_PREVIOUS, ia.GUESS_EXTRAPOLATE])
def test_graph_steps_equal_direct(precision, guess):
    a = ia.simple_spin(256, rel_tol=1e-9, precision=precision, guess=guess)
    b = ia.simple_spin(256, rel_tol=1e-9, precision=precision, guess=guess)
    a.set_launch(ia.LAUNCH_GRAPH)
    b.set_launch(ia.LAUNCH_DIRECT)
    dt = a.dt_reg
    for k in range(6):
        a.step_forward(dt)
        b.step_forward(dt)
        sa, sb = a.stats(), b.stats()
        assert sa == sb, (k, sa, sb)
    for nm in ("u", "v", "p", "phi", "X", "Y", "s"):
        assert np.array_equal(a.get(nm), b.get(nm)), nm
    a.close()
    b.close()


def test_graph_multibody_equal_direct():
    a = ia.multibody(256, 6, precision=ia.PREC_MIXED)
    b = ia.multibody(256, 6, precision=ia.PREC_MIXED)
    a.set_launch(ia.LAUNCH_GRAPH)
    b.set_launch(ia.LAUNCH_DIRECT)
    for k in range(5):
        a.step_forward(a.dt_reg)
        b.step_forward(b.dt_reg)
    for nm in ("u", "v", "p"):
        assert np.array_equal(a.get(nm), b.get(nm)), nm
    for k in range(a.nobj):
        assert np.array_equal(a.get("phi", k), b.get("phi", k)), k
    a.close()
    b.close()
