"""Many independent fluid_2d instances on one GPU (the batched-ensemble
configuration): every instance keeps its own HIP stream, and a pool of host
threads issues their steps (ctypes releases the interpreter lock during a
library call), so the kernels of small instances run side by side on the
GPU and one instance's host round trips do not stall the others.

Instances are independent: an instance stepped here gives bit for bit what
it gives stepped alone (every reduction is deterministic and every fused
smoother launch shape gives the same sweeps).
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor

from . import GUESS_PREVIOUS, PREC_FP64, simple_spin


class Ensemble:
    """specs: list of (m, omega) — an `ftest simple-spin m` rotor with
    angular-velocity prefactor omega (object.hh:422-467) per instance."""

    def __init__(self, specs, rel_tol=1e-8, precision=PREC_FP64, threads=16, frame=4 * math.pi / 240,
                 guess=GUESS_PREVIOUS):
        self.specs = list(specs)
        self.threads = threads
        self.members = [simple_spin(m, omega=om, rel_tol=rel_tol, precision=precision, guess=guess)
                        for m, om in self.specs]
        # the instances already run side by side, and a second stream per
        # instance (made by the first overlapped step) would share the
        # process's hardware queues with the streams that do the work: the
        # end-of-step rebuild stays in order
        for f in self.members:
            f.set_overlap(False)
        self.adt = [f.timestep_select(frame)[1] for f in self.members]
        # longest jobs first (largest grids), so the pool ends evenly
        self.order = sorted(range(len(self.members)), key=lambda i: -self.specs[i][0])

    def close(self):
        for f in self.members:
            f.close()
        self.members = []

    def steps(self, k):
        """k steps of every instance; returns per-instance wall seconds."""
        times = [0.0] * len(self.members)

        def run(i):
            t0 = time.perf_counter()
            self.members[i].steps(self.adt[i], k)
            self.members[i].sync()
            times[i] = time.perf_counter() - t0

        with ThreadPoolExecutor(self.threads) as ex:
            list(ex.map(run, self.order))
        return times

    def throughput(self, k):
        """Run k steps of every instance and report cell-steps/s: for the
        whole ensemble (total cells x k / wall time) and per grid size."""
        t0 = time.perf_counter()
        times = self.steps(k)
        wall = time.perf_counter() - t0
        per = {}
        for (m, _), t in zip(self.specs, times):
            e = per.setdefault(m, dict(instances=0, cell_steps=0, busy_s=0.0))
            e["instances"] += 1
            e["cell_steps"] += m * m * k
            e["busy_s"] += t
        total = sum(e["cell_steps"] for e in per.values())
        out = dict(wall_s=wall, cell_steps_per_s=total / wall, steps=k, instances=len(self.specs), sizes={})
        for m, e in sorted(per.items()):
            out["sizes"][m] = dict(instances=e["instances"], cell_steps=e["cell_steps"],
                                   cell_steps_per_s_in_ensemble=e["cell_steps"] / wall,
                                   mean_ms_per_step=1e3 * e["busy_s"] / (e["instances"] * k))
        return out
