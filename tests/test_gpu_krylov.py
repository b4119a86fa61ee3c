"""Mixed-precision solves as conjugate gradients preconditioned by one FP32
V-cycle per iteration (set_krylov): they reach the same relative residual
as V-cycle refinement, agree with its solutions to that tolerance, stay
deterministic (graph replay = direct launches, bit for bit), and fall back
to the refinement where the Krylov form is not built (periodic boxes)."""
import math

import numpy as np
import pytest

import incrmt_amd as ia

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,n", [(ia.MG_MAC, 256), (ia.MG_FEM, 257), (ia.MG_MAC, 192)])
def test_krylov_standalone_solve(kind, n):
    h = 2.0 / (n - 1 if kind == ia.MG_FEM else n)
    rng = np.random.default_rng(7)
    b = rng.standard_normal(n * n)
    b -= b.mean()
    sols = {}
    for kr in (False, True):
        s = ia.MgSolver(kind, n, n, h, h)
        s.set_krylov(kr)
        z, cyc, rel = s.solve(b, rel_tol=1e-11, precision=ia.PREC_MIXED)
        assert rel <= 1e-11, (kr, rel)
        # an independent residual: the solver's own operator applied on the host side
        r = b - s.apply(z)
        assert np.linalg.norm(r) <= 2e-11 * np.linalg.norm(b)
        sols[kr] = (z - z.mean(), cyc)
        s.close()
    (za, ca), (zb, cb) = sols[False], sols[True]
    assert np.abs(za - zb).max() <= 1e-8 * np.abs(za).max()
    assert cb <= ca


def test_krylov_steps_match_refinement_and_graphs():
    out = {}
    for kr, launch in ((False, ia.LAUNCH_DIRECT), (True, ia.LAUNCH_DIRECT), (True, ia.LAUNCH_GRAPH)):
        f = ia.simple_spin(128, omega=1.1, rel_tol=1e-11, precision=ia.PREC_MIXED, guess=ia.GUESS_QUADRATIC)
        f.set_krylov(kr)
        f.set_launch(launch)
        adt = f.timestep_select(4 * math.pi / 240)[1]
        f.steps(adt, 25)
        st = f.stats()
        assert st["mac_rel_res"] <= 1e-11 and st["fem_rel_res"] <= 1e-11
        out[(kr, launch)] = {nm: f.get(nm) for nm in ("u", "v", "p")}
        f.close()
    ref, kd, kg = out[(False, ia.LAUNCH_DIRECT)], out[(True, ia.LAUNCH_DIRECT)], out[(True, ia.LAUNCH_GRAPH)]
    for nm in ("u", "v", "p"):
        assert np.array_equal(kd[nm], kg[nm]), nm  # deterministic, graph or not
        scale = np.abs(ref[nm]).max()
        assert np.abs(kd[nm] - ref[nm]).max() <= 1e-7 * scale, nm


def test_krylov_periodic_keeps_refinement():
    a = ia.MgSolver(ia.MG_FEM, 64, 64, 1 / 64, 1 / 64, x_prd=True, y_prd=True)
    b = ia.MgSolver(ia.MG_FEM, 64, 64, 1 / 64, 1 / 64, x_prd=True, y_prd=True)
    b.set_krylov(True)
    rhs = np.sin(2 * np.pi * np.arange(64 * 64) / 64.0)
    rhs -= rhs.mean()
    za, ca, _ = a.solve(rhs, rel_tol=1e-10, precision=ia.PREC_MIXED)
    zb, cb, _ = b.solve(rhs, rel_tol=1e-10, precision=ia.PREC_MIXED)
    assert ca == cb and np.array_equal(za, zb)
    a.close()
    b.close()
