"""Oracle level set (oracle/levelset_oracle.cc) pinned against the reference.

The reference's level-set library (levelset/levelset.cc, heap.cc,
sorting.cc, traverse.hh) is compiled from its own sources into
oracle/_ref/libreflevel.so by oracle/Makefile. The oracle's restatement of
build_band and of the reference-map extension must agree with it bit for bit
on phi, on the band status codes and on the band traversal order. Closed
forms pin what the reference build would also get right: signed distance to
a circle and exact extension of linear fields.
"""
import numpy as np
import pytest

from tests import oracle_lib as ol
from tests import shapes

CASES = [("rotor", 0.0), ("rotor", 0.3), ("rotor", 2.1), ("circle", 0.0), ("wavy", 0.0), ("two", 0.0)]


def _args(m, limit_cells=7.0):
    geo = shapes.ls_geometry(m, m)
    lim = limit_cells * geo["dx"]
    return geo, (geo["m"], geo["n"], geo["ax"], geo["ay"], geo["dx"], geo["dy"], -lim, lim)


needs_ref = pytest.mark.skipif(not ol.have_reference(), reason="reference build absent (oracle/_ref)")


@needs_ref
@pytest.mark.parametrize("m", [32, 64, 128, 256])
@pytest.mark.parametrize("kind,ang", CASES)
def test_build_band_matches_reference(m, kind, ang):
    geo, args = _args(m)
    phi = shapes.shape(kind, geo, ang).ravel()
    po, so, oo = ol.build_band("oracle", *args, phi)
    pr, sr, orr = ol.build_band("reference", *args, phi)
    assert np.array_equal(so, sr)
    assert np.array_equal(po, pr)  # bitwise: same arithmetic, no FMA
    assert np.array_equal(oo, orr)


@needs_ref
@pytest.mark.parametrize("m", [64, 128])
def test_full_march_matches_reference(m):
    # ex_test's setting: cut-offs far away, so the march fills the grid
    geo, args = _args(m, limit_cells=1e6)
    phi = shapes.shape("wavy", geo).ravel()
    po, so, _ = ol.build_band("oracle", *args, phi)
    pr, sr, _ = ol.build_band("reference", *args, phi)
    assert np.array_equal(so, sr) and np.array_equal(po, pr)


@needs_ref
@pytest.mark.parametrize("m", [64, 128, 256])
@pytest.mark.parametrize("kind,ang", CASES)
def test_extension_matches_reference(m, kind, ang):
    geo, args = _args(m)
    X, Y = shapes.centres(geo)
    phi = shapes.shape(kind, geo, ang).ravel()
    inside = phi < 0
    fx = np.where(inside, ((X + 0.5 * Y) ** 2).ravel(), 0.0)
    fy = np.where(inside, (np.sin(3 * X) * Y).ravel(), 0.0)
    ro = ol.build_band_extrapolate("oracle", *args, phi, fx, fy)
    rr = ol.build_band_extrapolate("reference", *args, phi, fx, fy)
    assert ro[0] == rr[0] == 0
    for a, b in zip(ro[1:], rr[1:]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("m", [64, 128, 256])
def test_circle_band_is_signed_distance(m):
    # closed form: a circle's narrow band approximates |x-c|-R to O(dx^2)
    geo, args = _args(m)
    X, Y = shapes.centres(geo)
    exact = (np.sqrt((X - 0.1) ** 2 + (Y + 0.05) ** 2) - 0.5).ravel()
    phi, s, _ = ol.build_band("oracle", *args, exact.copy())
    band = (s >= 2) & (s <= 5)
    err = np.abs(phi[band] - exact[band]).max()
    assert err < 2.0 * geo["dx"] ** 2 / 0.5 + 1e-12
    interface = (s == 3) | (s == 4)
    assert np.abs(phi[interface] - exact[interface]).max() < 1e-3 * geo["dx"]
    # every point outside the band is clamped to the cut-off
    lim = 7 * geo["dx"]
    assert np.all(phi[s < 2] == -lim) and np.all(phi[s > 5] == lim)


@pytest.mark.parametrize("m", [64, 128])
def test_extension_reproduces_linear_fields(m):
    # a weighted least-squares linear fit reproduces linear data exactly
    geo, args = _args(m)
    X, Y = shapes.centres(geo)
    phi = shapes.shape("rotor", geo, 0.4).ravel()
    lin_x = (0.3 + 1.7 * X - 0.4 * Y).ravel()
    lin_y = (-0.2 + 0.5 * X + 2.5 * Y).ravel()
    fx = np.where(phi < 0, lin_x, 123.0)
    fy = np.where(phi < 0, lin_y, -77.0)
    rc, ph, s, ex, ey = ol.build_band_extrapolate("oracle", *args, phi, fx, fy)
    assert rc == 0
    band = (s == 4) | (s == 5)
    assert np.abs(ex[band] - lin_x[band]).max() < 1e-9
    assert np.abs(ey[band] - lin_y[band]).max() < 1e-9
