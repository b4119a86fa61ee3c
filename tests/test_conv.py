"""The resolution comparison of conv_test (fluid_2d::l_comparison,
fluid_2d_io.cc:186-281) as incrmt_amd.conv restates it, pinned by closed
forms on synthetic fields (no GPU): identical fields compare to zero, a
smooth field sampled at two resolutions compares to its interpolation
error, which falls with the stencil's order, and the point counts and
weights follow the reference's rules."""
import math

import numpy as np

from incrmt_amd import conv


SIN = lambda x, y: np.sin(x + 0.3 * y)
COS = lambda x, y: np.cos(0.7 * x - y)


def _fields(m, f_u, f_v, f_p, phis=(), eps=0.1):
    h = 2.0 / m
    c = -1 + h * (np.arange(m + 4) - 1.5)   # cell centres of the ghosted grid
    X, Y = np.meshgrid(c, c)
    Xn, Yn = X - 0.5 * h, Y - 0.5 * h        # lower-left corners
    return dict(m=m, n=m, u=f_u(X, Y), v=f_v(X, Y), p=f_p(Xn, Yn),
                phi=[ph(X, Y) for ph in phis], eps=eps)


def test_identical_fields_compare_to_zero():
    a = _fields(32, SIN, COS, lambda x, y: x * y)
    b = _fields(64, SIN, COS, lambda x, y: x * y)
    e = conv.l_comparison(a, b)
    # counts: all velocity points fluid (no solids), (m+1)^2 pressure corners
    assert e[0] == 1.0 and e[3] == 1.0
    assert math.isclose(e[7], 33 * 33 / (32 * 32))
    # the pressure at shared corners is the same function: zero difference
    assert e[13] == 0 and e[21] == 0 and e[31] == 0
    # velocity: cubic interpolation error of sin/cos at h = 1/32, small
    assert 0 < e[11] < 1e-4


def test_velocity_error_falls_fourth_order():
    # the even-ratio stencil is cubic interpolation: its error on a smooth
    # field falls as h^4
    f = lambda x, y: np.sin(2 * x) * np.cos(3 * y)
    errs = []
    for m in (16, 32, 64):
        a = _fields(m, f, f, lambda x, y: 0 * x)
        b = _fields(2 * m, f, f, lambda x, y: 0 * x)
        errs.append(conv.l_comparison(a, b)[11])
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(3.7 < r < 4.3 for r in rates), rates


def test_material_classes_and_boundary_weights():
    # a solid disc: velocity points split by min phi against eps
    disc = lambda x, y: np.sqrt(x * x + y * y) - 0.5
    a = _fields(32, SIN, COS, lambda x, y: 1 + 0 * x, phis=(disc,), eps=0.1)
    b = _fields(64, SIN, COS, lambda x, y: 2 + 0 * x, phis=(disc,), eps=0.1)
    e = conv.l_comparison(a, b)
    assert math.isclose(e[0] + e[1] + e[2], 1.0) and e[1] > 0 and e[2] > 0
    # pressure differs by 1 everywhere: L1 = sum of corner weights / (m n),
    # interior 1, edges 1/2, box corners 1/4 -> (m+1)^2 - 2(m+1) + 1 = m^2
    assert math.isclose(e[23], 1.0) and math.isclose(e[15], 1.0)
    assert e[31] == 1.0


def test_odd_ratio_uses_the_centre_value():
    f = lambda x, y: x + 2 * y
    a = _fields(16, f, f, lambda x, y: 0 * x)
    b = _fields(48, f, f, lambda x, y: 0 * x)
    # ratio 3: the fine cell at the coarse centre coincides with it
    assert conv.l_comparison(a, b)[11] < 1e-14


def test_periodic_corners_and_weights():
    """Along a periodic direction the pressure corners run to m-1 (n-1) and
    all weigh 1 (fluid_2d_io.cc:210-214)."""
    for px, py in ((True, False), (False, True), (True, True)):
        a = _fields(32, SIN, COS, lambda x, y: 1 + 0 * x)
        b = _fields(64, SIN, COS, lambda x, y: 3 + 0 * x)
        a.update(px=px, py=py)
        e = conv.l_comparison(a, b)
        cx, cy = (32 if px else 33), (32 if py else 33)
        assert math.isclose(e[7], cx * cy / (32 * 32)), (px, py)
        # the corner weights sum to m n either way (m+1 corners with halved
        # ends, or m whole ones): a constant difference 2 has L1 norm 2
        assert math.isclose(e[23], 2.0), (px, py, e[23])
        assert e[31] == 2
