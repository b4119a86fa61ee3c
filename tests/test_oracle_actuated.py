"""Actuated solids in the oracle: obj_flapper (object.hh:245-289,
object.cc:284-308) and obj_lamprey (object.hh:542-595, object.cc:554-580).
The reference's object.cc needs fluid_2d.hh and with it the absent TGMG
headers, so it cannot be built here; these functions are pinned instead by
what their definitions fix independently of the oracle's expressions:

- flapper phi is the distance to the segment |X| <= cl, Y = 0, less cr
  (computed here by projection onto the segment);
- lamprey phi vanishes at the nose (dis, 0), at the head-body joint
  (0, +-w1), along the tail end X = -la for |Y| <= w2, and is negative on
  the axis inside;
- actuation multiplies F's first column by s and the second by 1/s, so
  det F is preserved; for the flapper, s = exp(con Y) deep in the actuated
  rod with con = -log(m_stretch)/ar sin(2 pi t/T)^12 (evaluated here with
  numpy's power), s = 1 outside it and at t = 0; for the lamprey, at the
  centre of the contraction wave on the lower side (|Y| >= the local
  half-width) s = mlam once the ramp is over (t >= T/2) and sqrt(mlam) at
  t = T/4 (ramp o^2 (3 - 2o) = 1/2 at o = 1/2); the upper side is half a
  body length out of phase there (s = 1), and the head (X > eps) is never
  actuated."""
import ctypes
import math

import numpy as np
import pytest

from tests import oracle_lib as ol

D, I = ctypes.c_double, ctypes.c_int
EPS = 0.005


def ev(kind, p9, t, X, Y, F=(1.0, 0.0, 0.0, 1.0)):
    L = ol.oracle()
    L.orc_object_eval.argtypes = [I, ol._dp, D, D, D, D, ol._dp, ol._dp]
    p = np.array(p9, dtype=np.float64)
    phi = np.zeros(1)
    F4 = np.array(F, dtype=np.float64)
    L.orc_object_eval(kind, ol.dptr(p), EPS, t, X, Y, ol.dptr(phi), ol.dptr(F4))
    return phi[0], F4


FLAP = [0.0, -0.8, 0.026, 0.25, 0.021, 0.14, 0.0, 2.2, 8.0]   # ftest.cc:250
LAMP = [0.0, 0.0, 0.0, 0.8, 0.08, 0.04, 0.01, 1.0, 1.3]        # le 0.8, dis 0.08, w1 0.04, w2 0.01


def test_flapper_phi_is_segment_distance():
    rng = np.random.default_rng(0)
    cl, cr = FLAP[3], FLAP[2]
    for X, Y in rng.uniform(-0.4, 0.4, size=(200, 2)):
        want = math.hypot(X - min(max(X, -cl), cl), Y) - cr
        assert abs(ev(ol.FLAPPER, FLAP, 0.0, X, Y)[0] - want) <= 1e-15


@pytest.mark.parametrize("t", [1.3, 2.0, 5.7])
def test_flapper_actuation(t):
    ar, ms, T = FLAP[4], FLAP[7], FLAP[8]
    con = -math.log(ms) / ar * np.power(math.sin(2 * math.pi * t / T), 12)
    F = (1.3, -0.2, 0.4, 0.9)
    _, G = ev(ol.FLAPPER, FLAP, t, 0.05, 0.004, F)     # deep in the actuated rod: H = 1
    s = math.exp(con * 0.004)
    assert np.allclose(G, [F[0] * s, F[1] / s, F[2] * s, F[3] / s], rtol=1e-14, atol=0)
    assert abs((G[0] * G[3] - G[1] * G[2]) - (F[0] * F[3] - F[1] * F[2])) <= 1e-14
    _, G = ev(ol.FLAPPER, FLAP, t, 0.2, 0.004, F)      # outside the actuated region
    assert np.array_equal(G, F)
    _, G = ev(ol.FLAPPER, FLAP, 0.0, 0.05, 0.004, F)   # no contraction at t = 0
    assert np.array_equal(G, F)


def test_lamprey_phi_special_points():
    le, dis, w1, w2 = LAMP[3], LAMP[4], LAMP[5], LAMP[6]
    la = le - dis
    for X, Y in [(dis, 0.0), (0.0, w1), (0.0, -w1), (-la, 0.0), (-la, w2), (-la, -0.5 * w2)]:
        assert abs(ev(ol.LAMPREY, LAMP, 0.0, X, Y)[0]) <= 1e-15, (X, Y)
    assert ev(ol.LAMPREY, LAMP, 0.0, -0.5 * la, 0.0)[0] < 0
    assert ev(ol.LAMPREY, LAMP, 0.0, 0.5 * dis, 0.0)[0] < 0
    assert ev(ol.LAMPREY, LAMP, 0.0, -0.5 * la, 2 * w1)[0] > 0


def test_lamprey_actuation():
    le, dis, w1, w2, T, mlam = LAMP[3], LAMP[4], LAMP[5], LAMP[6], LAMP[7], LAMP[8]
    la = le - dis
    F = (1.1, 0.3, -0.2, 0.8)
    wl = lambda X: w1 - (w2 - w1) * X / la
    # t = T: shift la, wave centre of the lower side at X = -la/2
    X = -0.5 * la
    _, G = ev(ol.LAMPREY, LAMP, T, X, -2 * wl(X), F)
    assert np.allclose(G, [F[0] * mlam, F[1] / mlam, F[2] * mlam, F[3] / mlam], rtol=1e-14, atol=0)
    _, G = ev(ol.LAMPREY, LAMP, T, X, 2 * wl(X), F)        # upper side: out of phase
    assert np.array_equal(G, F)
    # t = T/4: ramp 1/2, shift la/4, wave centre at X = -3la/4
    X = -0.75 * la
    _, G = ev(ol.LAMPREY, LAMP, T / 4, X, -2 * wl(X), F)
    s = math.sqrt(mlam)
    assert np.allclose(G, [F[0] * s, F[1] / s, F[2] * s, F[3] / s], rtol=1e-14, atol=0)
    # the head is never actuated; det F is always preserved
    rng = np.random.default_rng(1)
    for t in (0.3, 0.9, 2.2):
        _, G = ev(ol.LAMPREY, LAMP, t, 0.5 * dis, 0.01, F)
        assert np.array_equal(G, F)
        for X, Y in rng.uniform(-0.8, 0.1, size=(20, 2)) * [1, 0.05]:
            _, G = ev(ol.LAMPREY, LAMP, t, X, Y, F)
            assert abs((G[0] * G[3] - G[1] * G[2]) - (F[0] * F[3] - F[1] * F[2])) <= 1e-13
