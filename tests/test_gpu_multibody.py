"""The sediment-style multibody configuration (many obj_circle solids,
collision stress between overlapping transition zones, fluid_2d.cc:455-500).

Per step against the oracle on a reduced case (8 circles from the same
placement rule at 128^2; stated tolerances as for the single-solid step:
u, v, p relative L2 <= 1e-6, X, Y within 1e-9, band status on >= 99.5% of
band points). At full size (2048^2 cells, 64 circles) the mixed-precision
run is compared every step with an FP64 run of the same case: both reach
the relative divergence 1e-8 in both projections, and the mixed run's
velocity stays within a relative L2 distance of 1e-6 of the FP64 run's.
"""
import numpy as np
import pytest

import incrmt_amd as ia
from incrmt_amd import workloads as wl
from tests import oracle_lib as ol
from tests.test_gpu_step import compare, gpu_from_spec, rel

pytestmark = pytest.mark.gpu


def spec_circles(m, count, seed=20260102):
    return dict(m=m, n=m, ax=-1.0, bx=1.0, ay=-1.0, by=1.0, visc=0.01, rhof=1.0, tw=2.5, pinw=2.5,
                w_scale=3.0, w_force=20.0, ev_mult=0.8, dt_pad=0.4, dt_ev_pad=0.8,
                objects=[(ol.CIRCLE, [cx, cy, r, 1.0, 0, 0], 1.0, 2.0, 1.0)
                         for cx, cy, r in wl.sediment_circles(count, seed, m, rmin=0.08, rmax=0.12)])


def test_eight_circles_per_step():
    m = 128
    spec = spec_circles(m, 8)
    g, o = gpu_from_spec(spec), ol.OracleFluid(spec, mg_tol=1e-13)
    dx = 2.0 / m
    for k in range(1, 6):
        g.step_forward(g.dt_reg)
        o.step(o.dt_reg)
        compare(g, o, m, dx, f"step {k}")


def test_2048_64_circles_mixed_vs_fp64():
    steps = 12
    a = ia.multibody(2048, 64, precision=ia.PREC_MIXED)
    b = ia.multibody(2048, 64, precision=ia.PREC_FP64)
    assert a.dt_reg == b.dt_reg
    inner = (slice(2, -2), slice(2, -2))
    for k in range(1, steps + 1):
        a.step_forward(a.dt_reg)
        b.step_forward(b.dt_reg)
        sa, sb = a.stats(), b.stats()
        for s in (sa, sb):
            assert s["mac_rel_res"] <= 1e-8 and s["fem_rel_res"] <= 1e-8
        if k % 4 == 0:
            du = rel(a.get("u")[inner], b.get("u")[inner])
            dv = rel(a.get("v")[inner], b.get("v")[inner])
            print(f"step {k}: mixed vs fp64 u {du:.2e} v {dv:.2e} cycles {sa['mac_cycles']}/{sa['fem_cycles']}",
                  flush=True)
            assert du <= 1e-6 and dv <= 1e-6


def _circles_fluid(spec, windows):
    f = ia.Fluid2D(spec["m"], spec["n"], visc=spec["visc"], rhof=spec["rhof"], tw=spec["tw"], pinw=spec["pinw"],
                   ax=spec["ax"], bx=spec["bx"], ay=spec["ay"], by=spec["by"])
    for kind, p, G, rhos, evm in spec["objects"]:
        f.add_circle(p[0], p[1], p[2], p[3], G=G, rhos=rhos, ev_trans_mult=evm)
    f.set_solver(rel_tol=1e-12, max_cycles=200)
    f.set_solid_windows(windows)
    f.initialize(spec["ev_mult"], spec["dt_pad"], spec["dt_ev_pad"])
    return f


def test_wall_cut_windows_bit_identical():
    """Windows of solids whose band box crosses the grid's edge: circles cut
    by one wall and by a corner, one just clear of a wall, one cut by two
    opposite walls (whole grid) and one in the middle; the windowed step
    equals the whole-grid step bit for bit over 8 steps."""
    spec = spec_circles(128, 1)
    spec["objects"] = [(ol.CIRCLE, [cx, cy, r, 1.0, 0, 0], 1.0, 2.0, 1.0) for cx, cy, r in [
        (-0.95, -0.3, 0.2),   # cut by the left wall
        (-1.1, 1.1, 0.45),    # cut by the left and top walls
        (-0.2, -0.8, 0.16),   # clear of the bottom wall by 3 cells
        (1.6, 0.0, 1.2),      # cut by the top and bottom walls
        (-0.3, 0.25, 0.15),   # in the middle
    ]]
    a, b = _circles_fluid(spec, True), _circles_fluid(spec, False)
    for k in range(1, 9):
        a.step_forward(a.dt_reg)
        b.step_forward(b.dt_reg)
    for nm in ("u", "v", "p"):
        assert np.array_equal(a.get(nm), b.get(nm)), nm
    for o in range(len(spec["objects"])):
        for nm in ("X", "Y", "phi", "s"):
            assert np.array_equal(a.get(nm, o), b.get(nm, o)), (nm, o)
    a.close()
    b.close()


@pytest.mark.parametrize("m,count", [(128, 8), (256, 12)])
def test_solid_windows_bit_identical(m, count):
    """Per-solid sweeps over each solid's band box widened by a few cells
    (the default) give every field of the whole-grid sweeps bit for bit:
    velocities, pressure, and every solid's reference map, level set and
    band status, over 8 steps (circles near a wall take a box clipped to
    the grid, the others a box)."""
    spec = spec_circles(m, count)
    a, b = _circles_fluid(spec, True), _circles_fluid(spec, False)
    for k in range(1, 9):
        a.step_forward(a.dt_reg)
        b.step_forward(b.dt_reg)
    for nm in ("u", "v", "p"):
        assert np.array_equal(a.get(nm), b.get(nm)), nm
    for o in range(count):
        for nm in ("X", "Y", "phi", "s"):
            assert np.array_equal(a.get(nm, o), b.get(nm, o)), (nm, o)
    a.close()
    b.close()
