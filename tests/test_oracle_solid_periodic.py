"""Solid periodic boundaries (fluid_2d::solid_prd_bc, fluid_2d.cc:784-835)
with obj_full in the oracle, pinned by a closed form: a rigid translation."""
import numpy as np

from tests import oracle_lib as ol


def test_translating_full_solid():
    """Closed form for solid periodic boundaries (fluid_2d.cc:784-835): an
    obj_full moving at a uniform velocity is a rigid translation, a steady
    state with X = x - U t. The stress stays zero only if the ghost layers
    continue the reference map across the seams with the right box-length
    shifts, so velocity and reference map stay exact to rounding."""
    spec = ol.conv_spec(32, 2, 280)
    spec["vel"] = 0
    o = ol.OracleFluid(spec, mg_tol=1e-12)
    u = o.get("u"); u[:] = 0.5; o.set("u", u)
    v = o.get("v"); v[:] = 0.0; o.set("v", v)
    X0 = o.get("X").copy()
    for k in range(5):
        o.step(0.01)
    u, v, X = o.get("u"), o.get("v"), o.get("X")
    assert np.abs(u[2:-2, 2:-2] - 0.5).max() <= 1e-14 and np.abs(v[2:-2, 2:-2]).max() <= 1e-14
    assert np.abs(X - (X0 - 0.5 * 0.05))[2:-2, 2:-2].max() <= 1e-14
    o.close()
