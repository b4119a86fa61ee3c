"""The end-of-step band rebuild of a lone solid on a second stream, beside
the FEM projection (Fluid2D.set_overlap): the fields after many steps equal,
bit for bit, those of the same run with the rebuild after the velocity
update on the step's stream (the reference's order, fluid_2d.cc:298-369)."""
import math

import numpy as np
import pytest

import incrmt_amd as ia

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,precision,guess,steps", [
    (128, ia.PREC_FP64, ia.GUESS_PREVIOUS, 60),
    (256, ia.PREC_MIXED, ia.GUESS_QUADRATIC, 30),
])
def test_overlapped_rebuild_bit_identical(m, precision, guess, steps):
    runs = []
    for ovl in (True, False):
        f = ia.simple_spin(m, omega=1.3, rel_tol=1e-10, precision=precision, guess=guess)
        f.set_overlap(ovl)
        adt = f.timestep_select(4 * math.pi / 240)[1]
        f.steps(adt, steps)
        runs.append({nm: f.get(nm) for nm in ("u", "v", "p", "X", "Y", "phi", "s")})
        runs[-1]["passes"] = f.stats()["reinit_passes"]
        f.close()
    a, b = runs
    for nm in ("u", "v", "p", "X", "Y", "phi", "s"):
        assert np.array_equal(a[nm], b[nm]), nm
    assert a["passes"] == b["passes"]
