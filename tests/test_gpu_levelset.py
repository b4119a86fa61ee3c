"""GPU narrow-band level set vs the oracle (itself bit-identical to the
reference's level-set library, tests/test_oracle_levelset.py).

Parity bar: on shapes without exact ties in phi the GPU band (phi and status)
and the extended fields are bit-identical to the oracle's. Where the input is
mirror-symmetric on the grid (the un-rotated rotor) the reference's serial
heap breaks exact ties in phi by its queue order; there the status codes
still match exactly and phi may differ on at most 3% of the band by at most
0.05 dx.
"""
import numpy as np
import pytest

import incrmt_amd as ia
from tests import oracle_lib as ol
from tests import shapes

pytestmark = pytest.mark.gpu

EXACT = [("rotor", 0.3), ("rotor", 0.7), ("rotor", 2.1), ("wavy", 0.0), ("circle", 0.0), ("two", 0.0)]


def _args(m, n=None, cells=7.0):
    geo = shapes.ls_geometry(m, n or m)
    lim = cells * geo["dx"]
    return geo, (geo["m"], geo["n"], geo["ax"], geo["ay"], geo["dx"], geo["dy"], -lim, lim)


@pytest.mark.parametrize("m", [32, 64, 128, 256, 512])
@pytest.mark.parametrize("kind,ang", EXACT)
def test_build_band_bit_exact(m, kind, ang):
    geo, args = _args(m)
    phi = shapes.shape(kind, geo, ang).ravel()
    po, so, _ = ol.build_band("oracle", *args, phi)
    pg, sg, passes = ia.build_band(*args, phi)
    assert np.array_equal(so, sg)
    assert np.array_equal(po, pg)
    assert passes < 200


@pytest.mark.parametrize("m", [64, 128, 256])
def test_build_band_symmetric_ties(m):
    geo, args = _args(m)
    phi = shapes.shape("rotor", geo, 0.0).ravel()
    po, so, _ = ol.build_band("oracle", *args, phi)
    pg, sg, _ = ia.build_band(*args, phi)
    assert np.array_equal(so, sg)
    band = (so >= 2) & (so <= 5)
    d = np.abs(po - pg)[band]
    assert (d > 0).mean() <= 0.03
    assert d.max() <= 0.05 * geo["dx"]


@pytest.mark.parametrize("m", [64, 128, 256])
@pytest.mark.parametrize("kind,ang", EXACT)
def test_extension_bit_exact(m, kind, ang):
    geo, args = _args(m)
    X, Y = shapes.centres(geo)
    phi = shapes.shape(kind, geo, ang).ravel()
    fx = np.where(phi < 0, ((X + 0.5 * Y) ** 2).ravel(), 0.0)
    fy = np.where(phi < 0, (np.sin(3 * X) * Y).ravel(), 0.0)
    rc, po, so, ex, ey = ol.build_band_extrapolate("oracle", *args, phi, fx, fy)
    pg, sg, gx, gy, passes = ia.extrapolate(*args, phi, fx, fy)
    assert rc == 0
    assert np.array_equal(so, sg) and np.array_equal(po, pg)
    assert np.array_equal(ex, gx) and np.array_equal(ey, gy)


def test_non_square_and_boundary_touching():
    # interface crossing the grid edge exercises the one-sided bicubic rows
    # and the Newton fail-safe (levelset.cc:137-195)
    geo, args = _args(80, 48)
    X, Y = shapes.centres(geo)
    phi = (np.sqrt((X + 0.93) ** 2 + (Y + 0.9) ** 2) - 0.35).ravel()
    po, so, _ = ol.build_band("oracle", *args, phi)
    pg, sg, _ = ia.build_band(*args, phi)
    assert np.array_equal(so, sg) and np.array_equal(po, pg)


def test_full_march():
    geo, args = _args(96, cells=1e6)
    phi = shapes.shape("wavy", geo).ravel()
    po, so, _ = ol.build_band("oracle", *args, phi)
    pg, sg, _ = ia.build_band(*args, phi)
    assert np.array_equal(so, sg) and np.array_equal(po, pg)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_no_interface(sign):
    geo, args = _args(32)
    phi = np.full(geo["m"] * geo["n"], 0.5 * sign)
    pg, sg, passes = ia.build_band(*args, phi)
    assert np.all(sg == (7 if sign > 0 else 0))
    assert np.all(pg == (args[7] if sign > 0 else args[6]))


def test_tiny_grid():
    geo, args = _args(8)
    X, Y = shapes.centres(geo)
    phi = (np.sqrt((X - 0.05) ** 2 + (Y + 0.1) ** 2) - 0.45).ravel()
    po, so, _ = ol.build_band("oracle", *args, phi)
    pg, sg, _ = ia.build_band(*args, phi)
    assert np.array_equal(so, sg) and np.array_equal(po, pg)
