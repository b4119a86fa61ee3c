"""Oracle step (oracle/fluid_oracle.cc) pinned by closed forms and invariants.

fluid_2d.cc itself cannot be compiled here (it needs the absent TGMG), so the
restatement of its sweeps is pinned by what the mathematics fixes: the
limiter's exactness and order, the timestep formula, a solid at rest staying
at rest, and both projections leaving a velocity whose discrete divergence
(the reference's own source-term formula applied to the projected field)
vanishes to the solver tolerance.
"""
import ctypes
import math

import numpy as np
import pytest

from tests import divergence as dv
from tests import oracle_lib as ol


def mono_diff(*f):
    L = ol.oracle()
    L.orc_mono_diff.restype = ctypes.c_double
    L.orc_mono_diff.argtypes = [ctypes.c_double] * 5
    return L.orc_mono_diff(*f)


def test_mono_diff_exact_on_linear_data():
    for a, b in ((0.3, 1.7), (-2.0, -0.25), (5.0, 1e-3)):
        f = [a + b * k for k in range(5)]
        assert mono_diff(*f) == pytest.approx(b, rel=1e-14)


def test_mono_diff_zero_at_extrema():
    assert mono_diff(0.0, 1.0, 2.0, 1.0, 0.0) == 0.0
    assert mono_diff(3.0, 1.0, 0.5, 2.0, 4.0) == 0.0


def test_mono_diff_fourth_order_on_smooth_data():
    # away from extrema the limiter is inactive and the slope is the
    # fourth-order centred difference (8(f1-f-1) - (f2-f-2))/12 per cell
    errs = []
    for h in (0.02, 0.01, 0.005):
        x0 = 0.3
        f = [math.sin(x0 + k * h) for k in (-2, -1, 0, 1, 2)]
        errs.append(abs(mono_diff(*f) / h - math.cos(x0)))
    assert 3.7 < math.log2(errs[0] / errs[1]) < 4.3
    assert 3.7 < math.log2(errs[1] / errs[2]) < 4.3


def test_timestep_closed_form():
    # simple-spin at 128: the extra-viscosity limit wins (fluid_2d.cc:152-192):
    # ex_visc = 0.8 h rhos sqrt(G/rhos); dt = 0.5 rhos/(2 ex_visc)/(2/h^2) 0.8
    f = ol.OracleFluid(ol.simple_spin_spec(128))
    h = 2.0 / 128
    ev = 0.8 * h * 3.0 * math.sqrt(48.0 / 3.0)
    assert f.dt_reg == pytest.approx(0.5 * 3.0 / (ev * 2) / (2 / h**2) * 0.8, rel=1e-14)
    assert f.dt_reg == pytest.approx(4.8828125e-4, rel=1e-12)
    l, adt = ol.timestep_select(4 * math.pi / 240, f.dt_reg)
    assert l == 108


def test_rotor_at_rest_stays_at_rest():
    # omega = 0: the anchor never turns, F = I gives zero elastic stress, so
    # the state is an equilibrium; only rounding may move it. At 128 cells the
    # wall-repulsion zone (w = 7.5 dx, fluid_2d.cc:1049) stays clear of the
    # arm tips at r = 0.8 plus the blur width.
    spec = ol.simple_spin_spec(128, omega=0.0)
    f = ol.OracleFluid(spec, mg_tol=1e-12)
    for _ in range(4):
        f.step(f.dt_reg)
    u, v = f.get("u")[2:-2, 2:-2], f.get("v")[2:-2, 2:-2]
    assert np.abs(u).max() < 1e-8 and np.abs(v).max() < 1e-8


@pytest.mark.parametrize("m", [48, 64])
def test_projections(m):
    spec = ol.simple_spin_spec(m)
    f = ol.OracleFluid(spec, mg_tol=1e-12)
    l, adt = ol.timestep_select(4 * math.pi / 240, f.dt_reg)
    for _ in range(3):
        f.step(adt)
    h = 2.0 / m
    # MAC is an exact projection: the corrected faces are divergence free
    # to the solver tolerance (fluid_2d.cc:1175-1236)
    dm = dv.mac_div(f.get("ul"), f.get("vd"), m, m, h, h)
    scale = np.abs(f.get("ul")).max() / h
    assert np.abs(dm).max() <= 1e-9 * scale
    # FEM is an approximate projection (exact bilinear stiffness for the
    # solve, a cell-centred gradient for the update, fluid_2d.cc:1297-1374):
    # it reduces the nodal divergence without removing it
    dstar = dv.fem_div(f.get("c0"), f.get("c1"), m, m, h, h)
    dnew = dv.fem_div(f.get("u"), f.get("v"), m, m, h, h)
    assert np.linalg.norm(dnew) < 0.5 * np.linalg.norm(dstar)
    # copy_pressure leaves a pressure with zero trapezoid-weighted mean
    # (fluid_2d.cc:1317-1348)
    p = f.get("p")[2:m + 3, 2:m + 3]
    w = np.ones((m + 1, m + 1))
    w[0, :] *= 0.5; w[-1, :] *= 0.5; w[:, 0] *= 0.5; w[:, -1] *= 0.5
    assert abs((w * p).sum()) <= 1e-12 * np.abs(p).sum()
