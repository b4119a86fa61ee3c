"""Pins the oracle's tracer support (fluid tracers and the rotor's track
tracer, fluid_2d.cc:1376-1484, object.cc:202-219, 313-417):

* the bicubic interpolation is bit-identical to the reference's own
  bi_interp.cc compiled from its sources (oracle/_ref/librefinterp.so), and
  reproduces polynomials it must reproduce exactly;
* the random positions follow glibc's rand() sequence from its default seed
  (the reference never seeds it), checked against the C library itself;
* the track tracer's correction lands on the anchored reference-map point,
  its records come every 10 steps, and its unwound angle follows the rotor.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib as ol


def test_glibc_rand_sequence_matches_the_c_library():
    libc = ctypes.CDLL(None)
    libc.srand(1)  # the default state of an unseeded program
    want = np.array([libc.rand() for _ in range(2000)], dtype=np.int32)
    assert np.array_equal(ol.glibc_rand_sequence(2000), want)


def _grid_fields(m, n, seed):
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, m)[None, :]
    y = np.linspace(0, 1, n)[:, None]
    fx = np.sin(3 * x + 1) * np.cos(2 * y) + 0.1 * rng.standard_normal((n, m))
    fy = np.cos(4 * x * y) + 0.1 * rng.standard_normal((n, m))
    return fx, fy


@pytest.mark.skipif(not ol.have_reference(), reason="reference sources not present")
@pytest.mark.parametrize("solid", [True, False])
def test_interp_bitwise_equal_to_reference(solid):
    import os
    if not os.path.exists(ol.REF_INTERP_SO):
        pytest.skip("reference interpolation not built")
    m, n = 13, 11
    ax, bx, ay, by = -1.15, 1.15, -1.1, 1.2
    fx, fy = _grid_fields(m, n, 1)
    rng = np.random.default_rng(2)
    # points in every kind of interval: first, last, bulk, and outside
    pts = list(zip(rng.uniform(ax - 0.2, bx + 0.2, 300), rng.uniform(ay - 0.2, by + 0.2, 300)))
    hx, hy = (bx - ax) / (m - 1), (by - ay) / (n - 1)
    pts += [(ax + 0.3 * hx, ay + 0.2 * hy), (bx - 0.3 * hx, by - 0.2 * hy), (ax + 0.5 * hx, by - 0.5 * hy),
            (bx - 0.1 * hx, ay + 0.9 * hy), (ax, ay), (bx, by)]
    for x, y in pts:
        a = ol.interp2("oracle", m, n, ax, bx, ay, by, fx, fy, x, y, solid)
        b = ol.interp2("reference", m, n, ax, bx, ay, by, fx, fy, x, y, solid)
        assert np.array_equal(a, b), (x, y, a, b)


def test_interp_reproduces_quadratics_and_interpolates():
    # the scheme is cubic convolution (Catmull-Rom) in the bulk and quadratic
    # in the first and last intervals: any quadratic field is reproduced
    # exactly everywhere, and arbitrary data are matched at the grid points
    m, n = 12, 10
    ax, bx, ay, by = 0.0, 1.1, 0.0, 0.9
    X, Y = np.meshgrid(np.linspace(ax, bx, m), np.linspace(ay, by, n))
    q = 0.3 + 0.7 * X - 0.2 * Y + 0.5 * X * X - 0.4 * X * Y + 0.9 * Y * Y
    c = np.random.default_rng(4).standard_normal((n, m))
    rng = np.random.default_rng(3)
    for x, y in zip(rng.uniform(ax, bx, 100), rng.uniform(ay, by, 100)):
        out = ol.interp2("oracle", m, n, ax, bx, ay, by, q, c, x, y)
        assert abs(out[0] - (0.3 + 0.7 * x - 0.2 * y + 0.5 * x * x - 0.4 * x * y + 0.9 * y * y)) < 1e-13
        assert abs(out[2] - (0.7 + x - 0.4 * y)) < 1e-12  # d/dx of the quadratic
        assert abs(out[4] - (-0.2 - 0.4 * x + 1.8 * y)) < 1e-12
    hx, hy = (bx - ax) / (m - 1), (by - ay) / (n - 1)
    for j in range(n):
        for i in range(m):
            out = ol.interp2("oracle", m, n, ax, bx, ay, by, q, c, ax + i * hx, ay + j * hy)
            assert abs(out[1] - c[j, i]) < 1e-12


def _spin_with_tracers(m, ntrace):
    f = ol.OracleFluid(ol.simple_spin_spec(m), mg_tol=1e-10)
    f.init_tracers(ntrace)
    return f


def test_tracers_start_outside_the_solid():
    m = 64
    f = _spin_with_tracers(m, 128)
    tr = f.tracers()
    assert tr.shape == (128, 4)
    assert np.all(np.abs(tr[:, :2]) <= 1)
    # every accepted position has bilinear phi >= eps (fluid_2d.cc:1393-1408):
    # in particular none lies in the rotor's body
    eps = 2.5 * 2.0 / m
    phi = f.get("phi")
    dx = 2.0 / m
    for x, y in tr[:, :2]:
        gx, gy = (x + 1) / dx - 0.5, (y + 1) / dx - 0.5
        i, j = int(gx), int(gy)
        i = min(max(i, -1), m - 1)
        j = min(max(j, -1), m - 1)
        fx, fy = gx - i, gy - j
        p = phi[j + 2:j + 4, i + 2:i + 4]
        v = (1 - fy) * (p[0, 0] * (1 - fx) + p[0, 1] * fx) + fy * (p[1, 0] * (1 - fx) + p[1, 1] * fx)
        assert v >= eps
    # the positions are the glibc sequence, drawn in pairs, rejecting some
    seq = ol.glibc_rand_sequence(2000)
    rnd = -1 + (2.0 / 2147483647.0) * seq.astype(np.float64)  # common.hh:28-30
    pairs = list(zip(rnd[0::2], rnd[1::2]))
    accepted = [p for p in pairs if p in set(map(tuple, tr[:, :2]))]
    assert len(accepted) == 128
    assert tuple(tr[0, :2]) == accepted[0]


def test_tracers_move_with_the_flow_and_records_every_ten_steps():
    m = 64
    f = _spin_with_tracers(m, 64)
    l, adt = ol.timestep_select(4 * math.pi / 240, f.dt_reg)
    f.frame_output()  # frame 0, as fluid_2d::solve writes before stepping
    tr0 = f.tracers().copy()
    for _ in range(25):
        f.step(adt)
    rec = f.track()
    # records at t = 0 (frame 0), then after steps 10 and 20
    assert rec.shape == (3, 4)
    assert rec[0, 0] == 0
    np.testing.assert_allclose(rec[1:, 0], [10 * adt, 20 * adt], rtol=1e-12)
    tr = f.tracers()
    assert np.all(np.abs(tr[:, :2]) <= 1)
    # a passive tracer moves by about dt*|u| per step
    assert np.max(np.abs(tr[:, :2] - tr0[:, :2])) < 25 * adt * 5
    assert np.max(np.abs(tr[:, :2] - tr0[:, :2])) > 0


def test_track_correction_lands_on_the_anchored_point():
    m = 64
    f = _spin_with_tracers(m, 0)
    l, adt = ol.timestep_select(4 * math.pi / 240, f.dt_reg)
    for _ in range(10):
        f.step(adt)
    st = f.track_state()
    assert st[9] == 1  # correction still on
    x, y, X0, Y0 = st[0], st[1], st[4], st[5]
    assert (X0, Y0) == (0.9 * 0.8, 0.0)
    # the corrected position maps back to (X0, Y0) through the bicubic
    # interpolation of the reference map (object.cc:383-417)
    ml = m + 4
    dx = 2.0 / m
    X = f.get("X")
    Y = f.get("Y")
    out = ol.interp2("oracle", ml, ml, -1 - 1.5 * dx, 1 + 1.5 * dx, -1 - 1.5 * dx, 1 + 1.5 * dx, X, Y, x, y)
    assert math.hypot(out[0] - X0, out[1] - Y0) < 1e-12


def test_track_angle_follows_the_rotor():
    # the arm tracer turns with the anchor angle pi(1 - cos t) (object.cc:175-181)
    m = 64
    f = _spin_with_tracers(m, 0)
    l, adt = ol.timestep_select(4 * math.pi / 240, f.dt_reg)
    for _ in range(4 * l):
        f.step(adt)
    rec = f.track()
    t, ang = rec[:, 0], rec[:, 3]
    theta = math.pi * (1 - np.cos(t))
    assert np.max(np.abs(ang - theta)) < 0.15
    assert np.all(np.diff(ang) >= -1e-3)  # monotone while the anchor turns forward
