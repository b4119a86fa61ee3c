"""Periodic boxes: GPU fluid_2d::step_forward with x_prd / y_prd against
the oracle, per step (set_boundaries' periodic copies, the Godunov seam of
edge_boundary_conditions, the periodic MAC correction, the periodic FEM
source, copy_pressure's wrapped node row and column and average_pressure's
weights; fluid_2d.cc:689-872, 1193-1235, 1317-1348), on conv_test's
initial field (sim_velocity_pulses, sim_type.cc:43-47) with and without
its circle, and with one periodic direction between two walls.

Stated tolerances, as for the no-slip box: u, v, p relative L2 <= 1e-6;
X, Y max |difference| <= 1e-9; band status equal on 99.5% of the band,
phi within 0.05 dx. The initial velocities agree to a few ulps (device
sin/cos/exp against the host's), which the tolerances absorb."""
import numpy as np
import pytest

import incrmt_amd as ia
from tests import oracle_lib as ol
from tests.test_gpu_step import compare

pytestmark = pytest.mark.gpu


def gpu_from_spec(spec, tol=1e-12):
    f = ia.Fluid2D(spec["m"], spec["n"], visc=spec["visc"], rhof=spec["rhof"], tw=spec["tw"], pinw=spec["pinw"],
                   ax=spec["ax"], bx=spec["bx"], ay=spec["ay"], by=spec["by"], x_prd=bool(spec.get("x_prd")),
                   y_prd=bool(spec.get("y_prd")), w_scale=spec["w_scale"], w_force=spec["w_force"])
    for k, ob in enumerate(spec["objects"]):
        kind, p, G, rhos, evm = ob[:5]
        if kind == ol.FULL:
            f.add_full(G=G, rhos=rhos, ev_trans_mult=evm)
        elif kind == ol.SPIN:
            f.add_spin(*p, G=G, rhos=rhos, ev_trans_mult=evm)
        elif kind == ol.SQUARE:
            f.add_square(p[0], p[1], p[2], p[3], G=G, rhos=rhos, ev_trans_mult=evm)
        else:
            f.add_circle(p[0], p[1], p[2], p[3], G=G, rhos=rhos, ev_trans_mult=evm)
        if len(ob) > 5:
            f.set_solid_velocity(k, ob[5])
    if spec.get("solid_prd_bc"):
        f.set_solid_periodic(True)
    f.set_sim_velocity(spec.get("vel", 0))
    f.set_solver(rel_tol=tol, max_cycles=200)
    f.initialize(spec["ev_mult"], spec["dt_pad"], spec["dt_ev_pad"])
    return f


def run_pair(spec, steps, dt, oracle_tol=1e-13):
    g, o = gpu_from_spec(spec), ol.OracleFluid(spec, mg_tol=oracle_tol)
    assert g.dt_reg == pytest.approx(o.dt_reg, rel=1e-15)
    dx = (spec["bx"] - spec["ax"]) / spec["m"]
    compare(g, o, spec["m"], dx, "init")
    for k in range(1, steps + 1):
        g.step_forward(dt)
        o.step(dt)
        compare(g, o, spec["m"], dx, f"step {k}")
    return g, o


def off_axis(spec):
    """conv_test's circle sits on the grid's horizontal mirror line, where
    the band rebuild meets exact phi ties that the reference's heap breaks
    by insertion order; the per-step comparisons move it off the mirror
    lines (as the no-slip circle tests do) so that both sides march the
    same band."""
    for o in spec["objects"]:
        o[1][0], o[1][1] = -0.0937, 0.0311
    return spec


@pytest.mark.parametrize("ctype", [0, 2, 3, 4, 5])
def test_conv_cases_per_step(ctype):
    """conv_test types 0 (fluid only), 2 (obj_full with solid periodic
    boundaries, fluid_2d.cc:784-835), 3 (square, obj_square), 4 (circle,
    density 3) and 5 (circle, density 1) in the doubly periodic box, 64^2,
    12 steps at a stable dt."""
    spec = off_axis(ol.conv_spec(64, ctype, 280))
    g, o = run_pair(spec, 12, 0.004)
    # the seam: ghost columns and rows repeat the opposite edge
    u = g.get("u")
    assert np.array_equal(u[2:-2, 0:2], u[2:-2, -4:-2]) and np.array_equal(u[2:-2, -2:], u[2:-2, 2:4])
    assert np.array_equal(u[0:2, :], u[-4:-2, :]) and np.array_equal(u[-2:, :], u[2:4, :])
    g.close()


@pytest.mark.parametrize("px,py", [(True, False), (False, True)])
def test_one_periodic_direction_per_step(px, py):
    """One periodic direction between two no-slip walls. (The four-roll
    field is not used for per-step parity: its exact zero crossings make the
    slope limiter's branches hinge on the last bit, so a one-ulp change of
    the initial velocity moves the oracle's own first step by ~4e-6.)"""
    spec = off_axis(ol.conv_spec(64, 4, 280))
    spec.update(x_prd=int(px), y_prd=int(py), vel=2)
    g, o = run_pair(spec, 10, 0.004)
    g.close()


def test_four_rolls_initial_field():
    """sim_four_rolls (sim_type.cc:30-35) as init_fields blends it with a
    circle's momentum (fluid_2d.cc:210-238), against the oracle."""
    spec = off_axis(ol.conv_spec(64, 4, 280))
    spec.update(vel=1)
    spec["objects"] = [ob[:5] for ob in spec["objects"]]  # the circle sets its own velocity
    g, o = gpu_from_spec(spec), ol.OracleFluid(spec, mg_tol=1e-13)
    for nm in ("u", "v"):
        assert np.abs(g.get(nm) - o.get(nm)).max() <= 4e-16, nm
    g.close()


def test_periodic_projection_properties():
    """128^2 pulses with the circle: after each step the MAC-projected
    faces are discretely divergence-free to the solver tolerance, and the
    periodic seam faces agree on both sides (ghost column m repeats cell 0's
    corrected left face)."""
    spec = ol.conv_spec(128, 4, 280)
    g = gpu_from_spec(spec, tol=1e-10)
    h = 2.0 / 128
    for k in range(5):
        g.step_forward(0.002)
        ul, vd = g.get("ul"), g.get("vd")
        div = (ul[2:-2, 3:-1] - ul[2:-2, 2:-2]) / h + (vd[3:-1, 2:-2] - vd[2:-2, 2:-2]) / h
        src = (g.get("u")[2:-2, 2:-2])
        assert np.abs(div).max() <= 1e-7 * max(1.0, np.abs(src).max() / h), k
        assert np.array_equal(ul[2:-2, -2], ul[2:-2, 2])
        assert np.array_equal(vd[-2, 2:-2], vd[2, 2:-2])
    g.close()


def test_unsupported_periodic_size():
    # one periodic direction of 80 cells halves to 5: refused (a doubly
    # periodic box of any size resamples instead)
    with pytest.raises(ia.IncrmtError) as e:
        f = ia.Fluid2D(80, 80, visc=1e-3, rhof=1.0, x_prd=True, y_prd=False)
        f.initialize(0.4, 0.4, 0.8)
    assert e.value.code == ia.ERR_UNSUPPORTED


@pytest.mark.parametrize("m", [280, 315, 360])
def test_conv_ladder_sizes_per_step(m):
    """conv_test's own ladder (mtab1 = 2520 ... 360, 315, 280,
    conv_test.cc:92-95): its three coarsest sizes, type 4 (circle of density
    3), 3 steps per size against the oracle. 280 and 360 halve to 35 and 45
    points and 315 is odd, so both projections coarsen through a resampled
    level (and 315's finest level is smoothed in nine colours). Both
    solvers to relative residual 1e-12 (the oracle's FEM solve of the first
    step's small source stalls a little above 1e-13 at 360)."""
    spec = off_axis(ol.conv_spec(m, 4, 280))
    g, o = run_pair(spec, 3, 0.001, oracle_tol=1e-12)
    g.close()


def test_square_in_walled_box_per_step():
    """obj_square (object.cc:41-44, 83-87) sedimenting in the no-slip box:
    its corners exercise the band rebuild's Newton on a non-smooth level
    set."""
    spec = ol.conv_spec(64, 3, 280)
    spec.update(x_prd=0, y_prd=0, vel=0, visc=0.01, dt_pad=0.4, dt_ev_pad=0.8, ev_mult=0.8)
    spec["objects"] = [(ol.SQUARE, [0.07, 0.11, 0.3, 1.0, 0.0, 0.0], 1.0, 2.0, 1.0)]
    g, o = gpu_from_spec(spec), ol.OracleFluid(spec, mg_tol=1e-13)
    dx = 2.0 / 64
    for k in range(1, 11):
        g.step_forward(g.dt_reg)
        o.step(o.dt_reg)
        compare(g, o, 64, dx, f"step {k}")
    g.close()


def test_full_solid_periodic_reference_map():
    """conv_test type 2 at 64^2: the reference map's ghost layers are its
    periodic image shifted by the box length (s_field::prd_bc), the level
    set holds no interface (obj_full's phi = -1 is cleaned to the band's
    lower cut-off, levelset.cc:408-413) and the solid fills every cell."""
    spec = ol.conv_spec(64, 2, 280)
    g = gpu_from_spec(spec)
    for k in range(4):
        g.step_forward(0.004)
    X, Y, s = g.get("X"), g.get("Y"), g.get("s")
    assert np.array_equal(X[2:-2, 0:2], X[2:-2, -4:-2] - 2.0) and np.array_equal(X[2:-2, -2:], X[2:-2, 2:4] + 2.0)
    assert np.array_equal(Y[0:2, :], Y[-4:-2, :] - 2.0) and np.array_equal(Y[-2:, :], Y[2:4, :] + 2.0)
    assert np.all(s == 0)
    phi = g.get("phi")
    assert np.all(phi == phi.flat[0]) and phi.flat[0] < -0.05
    g.close()


def test_full_solid_leaves_no_room_for_tracers():
    """fluid_2d::init_tracers draws tracer positions until they fall outside
    every solid (fluid_2d.cc:1376-1386), for ever when obj_full fills the
    box; the library stops after a bounded number of draws with an error."""
    f = ia.Fluid2D(32, 32, visc=1e-3, rhof=1.0, x_prd=True, y_prd=True)
    f.add_full(G=1.0, rhos=1.0)
    f.set_solid_periodic(True)
    f.set_ntrace(8)
    with pytest.raises(ia.IncrmtError) as e:
        f.initialize(0.4, 0.4, 0.8)
    assert e.value.code == ia.ERR_ARG
    f.close()


def test_translating_full_solid():
    """obj_full at a uniform velocity with solid periodic boundaries is a
    rigid translation (X = x - U t, zero stress): the GPU keeps it to
    rounding, as the oracle does (tests/test_oracle_solid_periodic.py)."""
    spec = ol.conv_spec(64, 2, 280)
    spec["vel"] = 0
    g = gpu_from_spec(spec)
    u = g.get("u"); u[:] = 0.5; g.set("u", u)
    v = g.get("v"); v[:] = 0.0; g.set("v", v)
    X0 = g.get("X").copy()
    for k in range(5):
        g.step_forward(0.005)
    u, v, X = g.get("u"), g.get("v"), g.get("X")
    assert np.abs(u[2:-2, 2:-2] - 0.5).max() <= 1e-13 and np.abs(v[2:-2, 2:-2]).max() <= 1e-13
    assert np.abs(X - (X0 - 0.5 * 0.025))[2:-2, 2:-2].max() <= 1e-13
    g.close()
