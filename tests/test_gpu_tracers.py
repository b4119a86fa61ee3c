"""GPU tracers vs the oracle (fluid_2d.cc:1376-1484, object.cc:202-219,
313-417): the 512 passive tracers of ftest and the rotor's track.dat.

Stated tolerances: initial tracer positions identical (both come from
glibc's rand() sequence and the same level set); after each step tracer
positions within 1e-9 (the bicubic velocity interpolation is the
reference's expression by expression; the velocity fields agree to the
step's tolerances); track records: times identical, positions within 1e-9,
unwound angle within 1e-8.

The long run is the `ftest simple-spin 128` case to t = 2 pi (frame 120,
12,960 steps) against the oracle's run of the same case saved in
tests/golden/spin128_frames.npz (made by tests/golden/make_spin128_frames.py):
u, v relative L2 <= 1e-5, phi and X, Y within 1e-6 inside the band, every
track record within 1e-6 and the tracers within 1e-5.
"""
import math
import os

import numpy as np
import pytest

import incrmt_amd as ia
from tests import oracle_lib as ol
from tests.test_gpu_step import rel

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "spin128_frames.npz")


def _pair(m, ntrace, tol=1e-12):
    spec = ol.simple_spin_spec(m)
    o = ol.OracleFluid(spec, mg_tol=1e-13)
    o.init_tracers(ntrace)
    f = ia.Fluid2D(m, m, visc=spec["visc"], rhof=spec["rhof"], tw=spec["tw"], pinw=spec["pinw"])
    for kind, p, G, rhos, evm in spec["objects"]:
        f.add_spin(*p, G=G, rhos=rhos, ev_trans_mult=evm)
    f.set_solver(rel_tol=tol, max_cycles=200)
    f.set_ntrace(ntrace)
    f.initialize(spec["ev_mult"], spec["dt_pad"], spec["dt_ev_pad"])
    return f, o


def test_initial_tracers_identical():
    f, o = _pair(64, 512)
    assert np.array_equal(f.tracers()[:, :2], o.tracers()[:, :2])
    assert np.array_equal(f.track_state(), o.track_state())


def test_tracers_and_track_per_step():
    m = 64
    f, o = _pair(m, 256)
    l, adt = f.timestep_select(4 * math.pi / 240)
    f.frame_output()
    o.frame_output()
    for k in range(1, 36):
        f.step_forward(adt)
        o.step(adt)
        d = np.abs(f.tracers()[:, :2] - o.tracers()[:, :2]).max()
        assert d <= 1e-9, f"step {k}: tracer diff {d:.3e}"
        st_g, st_o = f.track_state(), o.track_state()
        assert np.abs(st_g[:4] - st_o[:4]).max() <= 1e-9, f"step {k}: track {st_g} {st_o}"
        assert np.array_equal(st_g[4:], st_o[4:]), f"step {k}"
    f.frame_output()
    o.frame_output()
    rg, ro = f.track(), o.track()
    assert rg.shape == ro.shape == (5, 4)  # frame 0, steps 10, 20, 30, frame end
    assert np.array_equal(rg[:, 0], ro[:, 0])
    assert np.abs(rg[:, 1:3] - ro[:, 1:3]).max() <= 1e-9
    assert np.abs(rg[:, 3] - ro[:, 3]).max() <= 1e-8


def test_tracers_leaving_the_box_are_remapped():
    # tracer_remap (sim_type.cc:12-17): a tracer pushed out through a wall
    # re-enters on the other side with a fresh rand() coordinate; the GPU
    # draws the same numbers in the same order as the oracle
    m = 64
    f, o = _pair(m, 16)
    tr = o.tracers()
    l, adt = f.timestep_select(4 * math.pi / 240)
    # two tracers placed outside the box: both sides move them with the
    # extrapolated velocity and remap them at the end of the step
    tr[3, 0] = 1.5
    tr[7, 1] = -1.5
    f.set_tracers(tr)
    o.set_tracers(tr)
    f.step_forward(adt)
    o.step(adt)
    a, b = f.tracers(), o.tracers()
    assert np.abs(a[:, :2] - b[:, :2]).max() <= 1e-9
    assert np.all(np.abs(a[:, :2]) <= 1)


@pytest.mark.slow
@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="golden frames not generated")
def test_spin128_to_two_pi_against_golden():
    g = np.load(GOLDEN)
    m = 128
    f, _ = _pair(m, 512)
    l, adt = f.timestep_select(4 * math.pi / 240)
    assert l == int(g["l"]) and adt == float(g["adt"])
    f.frame_output()
    inner = (slice(2, -2), slice(2, -2))
    for k in range(1, 121):
        f.steps(adt, l)
        f.frame_output()
        if k % 20 == 0:
            print(f"frame {k} t={f.time:.6f}", flush=True)
        if k in (60, 120):
            assert f.time == pytest.approx(float(g[f"time_{k}"]), rel=1e-13)
            for nm in ("u", "v"):
                r = rel(f.get(nm)[inner], g[f"{nm}_{k}"][inner])
                assert r <= 1e-5, f"frame {k}: {nm} relative L2 {r:.3e}"
            phi_o = g[f"phi_{k}"]
            band = np.abs(phi_o) < 0.05
            for nm in ("phi", "X", "Y"):
                d = np.abs(f.get(nm) - g[f"{nm}_{k}"])[band].max()
                assert d <= 1e-6, f"frame {k}: {nm} max diff in band {d:.3e}"
            d = np.abs(f.tracers()[:, :2] - g[f"tracers_{k}"]).max()
            assert d <= 1e-5, f"frame {k}: tracer diff {d:.3e}"
    rec = f.track()
    want = g["track"]
    assert rec.shape == want.shape
    assert np.abs(rec[:, 0] - want[:, 0]).max() <= 1e-12
    assert np.abs(rec[:, 1:] - want[:, 1:]).max() <= 1e-6
