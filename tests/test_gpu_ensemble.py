"""The batched-ensemble configuration (independent rotor instances of
uneven size stepped concurrently, incrmt_amd/ensemble.py). Stated bar:
every instance stepped in the ensemble equals, bit for bit, the same
instance stepped alone (u, v, p, X, Y, band status)."""
import numpy as np
import pytest

import incrmt_amd as ia
from incrmt_amd.ensemble import Ensemble
from incrmt_amd.workloads import rotor_ensemble

pytestmark = pytest.mark.gpu


def test_ensemble_members_equal_solo_runs():
    specs = rotor_ensemble(10, sizes=(64, 128))
    e = Ensemble(specs, rel_tol=1e-10, threads=4)
    e.steps(4)
    for (m, om), f, adt in zip(specs, e.members, e.adt):
        solo = ia.simple_spin(m, omega=om, rel_tol=1e-10)
        solo.steps(adt, 4)
        for nm in ("u", "v", "p", "X", "Y", "s"):
            assert np.array_equal(f.get(nm), solo.get(nm)), (m, om, nm)
        solo.close()
    e.close()


def test_ensemble_mixed_quadratic_equal_solo_runs():
    """Mixed precision with quadratic starts: the first solves of some
    instances set up their start buffers while other host threads capture
    V-cycle graphs; every instance still equals its solo run."""
    specs = rotor_ensemble(12, sizes=(64, 128))
    kw = dict(rel_tol=1e-10, precision=ia.PREC_MIXED, guess=ia.GUESS_QUADRATIC)
    e = Ensemble(specs, threads=6, **kw)
    e.steps(4)
    for (m, om), f, adt in zip(specs, e.members, e.adt):
        solo = ia.simple_spin(m, omega=om, **kw)
        solo.steps(adt, 4)
        for nm in ("u", "v", "p", "X", "Y", "s"):
            assert np.array_equal(f.get(nm), solo.get(nm)), (m, om, nm)
        solo.close()
    e.close()


def test_ensemble_256_members_equal_solo_runs():
    """The full configuration: 256 rotors, sizes 64..512 drawn by seed
    20260103, the headline solver, 32 host threads, three steps; every
    member equals its solo run bit for bit."""
    specs = rotor_ensemble(256)
    kw = dict(rel_tol=1e-8, precision=ia.PREC_MIXED, guess=ia.GUESS_QUADRATIC)
    e = Ensemble(specs, threads=32, **kw)
    e.steps(3)
    bad = []
    for (m, om), f, adt in zip(specs, e.members, e.adt):
        solo = ia.simple_spin(m, omega=om, **kw)
        solo.steps(adt, 3)
        bad += [(m, om, nm) for nm in ("u", "v", "p", "X", "Y", "s") if not np.array_equal(f.get(nm), solo.get(nm))]
        solo.close()
    e.close()
    assert sorted({m for m, _ in specs}) == [64, 128, 256, 512]
    assert not bad, bad[:8]


def test_ensemble_throughput_reports_every_size():
    specs = rotor_ensemble(24)
    e = Ensemble(specs, rel_tol=1e-8, threads=8)
    e.steps(1)
    r = e.throughput(3)
    print({k: (v["instances"], round(v["mean_ms_per_step"], 2)) for k, v in r["sizes"].items()},
          f"{r['cell_steps_per_s']:.3e} cell-steps/s", flush=True)
    assert sum(v["instances"] for v in r["sizes"].values()) == 24
    assert r["cell_steps_per_s"] > 0
    e.close()
