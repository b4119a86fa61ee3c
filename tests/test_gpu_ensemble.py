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
