"""The C++ driver written on the reference-side adapters
(include/incrmt_adapter.hh) runs `ftest simple-spin 64` and an mgs_base
solve through the library, and gives bit for bit what the Python binding
gives for the same calls."""
import math
import os
import subprocess

import numpy as np
import pytest

import incrmt_amd as ia

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "incrmt_amd", "adapter_driver")


def test_cxx_driver_matches_binding(tmp_path):
    m, steps = 64, 12
    out = tmp_path / "out.bin"
    print("adapter driver", m, steps, flush=True)
    r = subprocess.run([DRIVER, str(m), str(steps), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = np.fromfile(out, dtype=np.float64)
    N = (m + 4) ** 2
    f = ia.simple_spin(m, rel_tol=1e-8)
    l, adt = f.timestep_select(4 * math.pi / 240)
    f.steps(adt, steps)
    for k, nm in enumerate(("u", "v", "p")):
        assert np.array_equal(d[k * N:(k + 1) * N], f.get(nm).ravel()), nm
    f.close()
    h = 2.0 / m
    src = d[3 * N:3 * N + m * m]  # the source the driver formed
    assert abs(src.sum()) < 1e-9 * np.abs(src).sum()
    s = ia.MgSolver(ia.MG_MAC, m, m, h, h)
    z, cyc, rr = s.solve(src, rel_tol=1e-10)
    assert rr <= 1e-10 and np.array_equal(d[3 * N + m * m:], z)
    s.close()
