"""Actuated solids on the GPU against the oracle, per step: obj_flapper
(object.hh:245-289, object.cc:284-308; ftest's flapper geometry,
ftest.cc:244-251, with a short period so that the contraction is strong
within the steps compared) and obj_lamprey (object.hh:542-595,
object.cc:554-580), in no-slip boxes. The actuation enters the solid
stress at every edge (fluid_2d.cc:533-541) with pre_stress_setup's state at
the stress evaluation's time (initialisation, the half step and the end of
the step). Tolerances as for the other solids: u, v, p relative L2 <= 1e-6;
X, Y within 1e-9; band status equal on 99.5% of the band, phi within
0.05 dx."""
import numpy as np
import pytest

import incrmt_amd as ia
from tests import oracle_lib as ol
from tests.test_gpu_step import compare

pytestmark = pytest.mark.gpu


def spec_for(kind, m):
    if kind == ol.FLAPPER:
        objs = [(ol.FLAPPER, [0.0, -0.8, 0.026, 0.25, 0.021, 0.14, 0.3, 2.2, 0.05], 10.0, 4.0, 1.0)]
        box, visc = 1.5, 5e-4
    else:
        objs = [(ol.LAMPREY, [0.013, -0.021, 0.3, 0.8, 0.08, 0.04, 0.01, 0.05, 1.3], 10.0, 1.5, 1.0)]
        box, visc = 1.0, 1e-3
    return dict(m=m, n=m, ax=-box, bx=box, ay=-box, by=box, visc=visc, rhof=1.0, tw=2.5, pinw=2.5,
                w_scale=3.0, w_force=20.0, ev_mult=0.8, dt_pad=0.4, dt_ev_pad=0.8, objects=objs)


def gpu_from_spec(spec, tol=1e-12):
    f = ia.Fluid2D(spec["m"], spec["n"], visc=spec["visc"], rhof=spec["rhof"], tw=spec["tw"], pinw=spec["pinw"],
                   ax=spec["ax"], bx=spec["bx"], ay=spec["ay"], by=spec["by"])
    for kind, p, G, rhos, evm in spec["objects"]:
        if kind == ol.FLAPPER:
            f.add_flapper(*p, G=G, rhos=rhos, ev_trans_mult=evm)
        else:
            f.add_lamprey(*p, G=G, rhos=rhos, ev_trans_mult=evm)
    f.set_solver(rel_tol=tol, max_cycles=200)
    f.initialize(spec["ev_mult"], spec["dt_pad"], spec["dt_ev_pad"])
    return f


@pytest.mark.parametrize("kind,m", [(ol.FLAPPER, 128), (ol.LAMPREY, 128)])
def test_actuated_per_step(kind, m):
    spec = spec_for(kind, m)
    g, o = gpu_from_spec(spec), ol.OracleFluid(spec, mg_tol=1e-13)
    assert g.dt_reg == pytest.approx(o.dt_reg, rel=1e-15)
    dx = (spec["bx"] - spec["ax"]) / m
    compare(g, o, m, dx, "init")
    dt = min(g.dt_reg, 0.002)
    for k in range(1, 13):
        g.step_forward(dt)
        o.step(dt)
        compare(g, o, m, dx, f"step {k}")
    g.close()
    o.close()


def test_flapper_contracts():
    """Over one period the actuated flapper bends: its reference map
    departs from the rigid rotation of its initial state by more than the
    passive case's (same solid with m_stretch = 1, no actuation)."""
    out = []
    for ms in (2.2, 1.0):
        spec = spec_for(ol.FLAPPER, 128)
        spec["objects"][0][1][7] = ms
        g = gpu_from_spec(spec, tol=1e-8)
        X0 = g.get("X").copy()
        dt = min(g.dt_reg, 0.002)
        for k in range(25):
            g.step_forward(dt)
        band = (g.get("phi") < 0)
        out.append(np.abs(g.get("X") - X0)[band].max())
        g.close()
    assert out[0] > 2 * out[1], out
