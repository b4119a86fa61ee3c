"""Writes tests/golden/spin128_frames.npz: the simple-spin rotor of
`ftest simple-spin 128` (ftest.cc:74-143) run by the oracle as
fluid_2d::solve(4*pi, 240) runs it (fluid_2d.cc:260-294) up to frame 120
(t = 2 pi, 120 x 108 steps), with 512 tracers and the rotor's track tracer.

Saved: at frames 60 and 120 the velocity u, v, the level set phi and the
reference map X, Y on the ghosted grid and the 512 tracer positions, as
float32 (their rounding, below 1e-7, is far inside the comparison's
tolerances); the times and the whole track.dat record list as float64. Both projections are solved to a
relative residual of 1e-12, so the solver choice is invisible at the
tolerances the comparison uses.

Run from the repository root after building the oracle:
    python tests/golden/make_spin128_frames.py
It takes about 6 minutes on 8 cores.
"""
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402

M = 128
FRAMES = 240
DURATION = 4 * math.pi
STOP = 120


def main():
    f = ol.OracleFluid(ol.simple_spin_spec(M), mg_tol=1e-12)
    f.init_tracers(512)
    l, adt = ol.timestep_select(DURATION / FRAMES, f.dt_reg)
    f.frame_output()  # frame 0
    out = {"l": l, "adt": adt, "dt_reg": f.dt_reg}
    t0 = time.time()
    for k in range(1, STOP + 1):
        for _ in range(l):
            f.step(adt)
        f.frame_output()
        if k in (60, 120):
            for name in ("u", "v", "phi", "X", "Y"):
                out[f"{name}_{k}"] = f.get(name).astype(np.float32)
            out[f"tracers_{k}"] = f.tracers()[:, :2].astype(np.float32)
            out[f"time_{k}"] = f.time
        if k % 10 == 0:
            print(f"frame {k} t={f.time:.6f} {time.time() - t0:.0f}s", flush=True)
    out["track"] = f.track()
    np.savez_compressed(os.path.join(HERE, "spin128_frames.npz"), **out)


if __name__ == "__main__":
    main()
