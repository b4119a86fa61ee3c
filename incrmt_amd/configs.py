"""Runs of the benchmark configurations other than the headline step
(bench.py measures that one), printing one JSON line each.

    python -m incrmt_amd.configs fem        standalone FEM projection, 1024^2 nodes
    python -m incrmt_amd.configs multibody  2048^2 cells, 64 circles, mixed vs FP64
    python -m incrmt_amd.configs ensemble   256 rotor instances of uneven size
    python -m incrmt_amd.configs conv       resolution ladder of the rotor (conv_test)
    python -m incrmt_amd.configs conv-test  conv_test itself: periodic pulses, types 0..7
"""
import argparse
import json
import math
import sys
import time

import numpy as np

import incrmt_amd as ia
from incrmt_amd import workloads as wl


# Algorithmic bytes of one V-cycle per fine node, FP64: the four smoothing
# sweeps (z, b and the cell weight read, z written: 32 B each), the residual
# and restriction (z, b, weight read: 24 B) and the cycle's residual norm
# (24 B), the coarser levels adding a third (a quarter per level). In FP32
# refinement the sweeps and transfers move half of that, and the FP64
# residual, correction and norm 24 + 16 + 24 B.
FEM_VCYCLE_BYTES = {"fp64": (4 * 32 + 24 + 24) * 4 / 3, "mixed": (4 * 16 + 12) * 4 / 3 + 24 + 16 + 24}


def run_fem(nodes=1024, tol=1e-10, repeats=3):
    """Seeded Fourier right-hand side (workloads.fourier_fem_rhs), zero start,
    relative residual tol; FP64 V-cycles vs FP32 V-cycles in FP64 iterative
    refinement. The solve is timed on the device with b resident (HIP
    events around the solve alone, no host copy inside), the median of
    `repeats` solves; per V-cycle against the algorithmic bytes of a cycle
    (FEM_VCYCLE_BYTES) at the 8 TB/s HBM peak."""
    b, h = wl.fourier_fem_rhs(nodes)
    b = b.ravel() - b.mean()
    out = dict(config="standalone FEM projection", nodes=[nodes, nodes], rel_tol=tol,
               timing="device (HIP events around the solve, b resident, zero start)")
    for name, prec in (("fp64", ia.PREC_FP64), ("mixed", ia.PREC_MIXED)):
        s = ia.MgSolver(ia.MG_FEM, nodes, nodes, h, h)
        s.solve(b, rel_tol=tol, max_cycles=100, precision=prec)  # uploads b, warms up
        t = [s.time_solve(tol, 100, prec, reps=1) for _ in range(repeats)]
        ms = float(np.median([x[0] for x in t]))
        cyc, rr = t[0][1], t[0][2]
        per = ms / cyc
        alg = FEM_VCYCLE_BYTES[name] * nodes * nodes
        out[name] = dict(v_cycles=cyc, rel_residual=rr, ms_per_solve=ms, ms_per_v_cycle=per,
                         algorithmic_bytes_per_v_cycle=alg, algorithmic_ms_at_peak=alg / 8e12 * 1e3,
                         v_cycle_vs_algorithmic_time=per / (alg / 8e12 * 1e3))
        s.close()
    out["mixed_speedup_per_solve"] = out["fp64"]["ms_per_solve"] / out["mixed"]["ms_per_solve"]
    return out


def run_multibody(m=2048, count=64, steps=200, every=1):
    """The mixed-precision multibody run against an FP64 run of the same case,
    every `every` steps: relative divergence of both projections and the
    relative L2 distance of the mixed run's velocity from the FP64 run's."""
    a = ia.multibody(m, count, precision=ia.PREC_MIXED)
    b = ia.multibody(m, count, precision=ia.PREC_FP64)
    inner = (slice(2, -2), slice(2, -2))
    rows = []
    ta = tb = 0.0
    for k in range(1, steps + 1):
        t0 = time.perf_counter()
        a.step_forward(a.dt_reg)
        a.sync()
        t1 = time.perf_counter()
        b.step_forward(b.dt_reg)
        b.sync()
        t2 = time.perf_counter()
        ta += t1 - t0
        tb += t2 - t1
        if k % every == 0 or k == steps:
            sa, sb = a.stats(), b.stats()
            ua, ub, va, vb = a.get("u")[inner], b.get("u")[inner], a.get("v")[inner], b.get("v")[inner]
            err = math.sqrt((np.sum((ua - ub) ** 2) + np.sum((va - vb) ** 2)) / (np.sum(ub ** 2) + np.sum(vb ** 2)))
            rows.append(dict(step=k, mac_rel_div=sa["mac_rel_res"], fem_rel_div=sa["fem_rel_res"],
                             fp64_mac_rel_div=sb["mac_rel_res"], fp64_fem_rel_div=sb["fem_rel_res"],
                             rel_l2_vs_fp64=err, mac_cycles=sa["mac_cycles"], fem_cycles=sa["fem_cycles"]))
            print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
    out = dict(config="multibody", cells=[m, m], solids=count, steps=steps,
               mixed_ms_per_step=1e3 * ta / steps, fp64_ms_per_step=1e3 * tb / steps,
               mixed_cell_steps_per_s=m * m * steps / ta, fp64_cell_steps_per_s=m * m * steps / tb,
               max_rel_l2_vs_fp64=max(r["rel_l2_vs_fp64"] for r in rows),
               max_rel_div=max(max(r["mac_rel_div"], r["fem_rel_div"]) for r in rows), per_step=rows)
    a.close()
    b.close()
    return out


def run_ensemble(count=256, steps=10, threads=32, precision=ia.PREC_MIXED, guess=ia.GUESS_QUADRATIC):
    """256 simple-spin rotors of sizes drawn by seed 20260103 from
    64/128/256/512, stepped concurrently (one stream per instance). The
    default is the headline solver (mixed precision, quadratic starts) on
    32 host threads; precision=PREC_FP64, guess=GUESS_PREVIOUS gives the
    reference's own solver."""
    from incrmt_amd.ensemble import Ensemble
    specs = wl.rotor_ensemble(count)
    e = Ensemble(specs, rel_tol=1e-8, precision=precision, threads=threads, guess=guess)
    e.steps(2)  # the quadratic start needs two earlier solutions
    r = e.throughput(steps)
    r.update(config="batched ensemble", threads=threads, precision="fp64" if precision == ia.PREC_FP64 else "mixed",
             guess={ia.GUESS_PREVIOUS: "previous", ia.GUESS_EXTRAPOLATE: "extrapolate",
                    ia.GUESS_QUADRATIC: "quadratic", ia.GUESS_LEARNED: "learned"}[guess])
    e.close()
    return r


def run_conv(sizes=(128, 256, 512, 1024), duration=0.2):
    """conv_test's procedure (conv_test.cc:100-194) on the simple-spin rotor
    in its no-slip box: each size run to `duration` and compared with the
    finest by l_comparison; observed orders of the total velocity L2 and
    pressure L2 errors between successive sizes."""
    from incrmt_amd import conv
    t0 = time.perf_counter()
    errs, steps = conv.ladder(lambda m: ia.simple_spin(m, rel_tol=1e-10), sizes, duration)
    out = dict(config="conv ladder, simple-spin rotor", sizes=list(sizes), duration=duration, steps=steps,
               errors={m: [float(x) for x in e] for m, e in errs.items()},
               velocity_l2_orders=conv.orders(errs, 11), pressure_l2_orders=conv.orders(errs, 15),
               seconds=time.perf_counter() - t0)
    return out


def run_conv_test(ctype=4, sizes=(1536, 768, 512, 384, 256), intervals=50, every=10):
    """conv_test.cc:14-194 for one test type on a ladder of periodic sizes
    that all divide the finest (the reference's ladder 2520..280 has odd
    factors that this solver does not coarsen; 2^k and 3 * 2^k sizes do):
    the error table every `every` intervals, and the observed orders of
    the total velocity and pressure L2 errors at the last interval."""
    from incrmt_amd import conv
    t0 = time.perf_counter()

    def cb(i, errs):
        if i % every == 0 or i == intervals:
            print(json.dumps(dict(interval=i, vel_l2={m: e[11] for m, e in errs.items()},
                                  p_l2={m: e[15] for m, e in errs.items()})), file=sys.stderr, flush=True)

    table, steps = conv.conv_test_run(ctype, sizes, intervals=intervals, callback=cb)
    last = table[-1]
    return dict(config="conv_test", test_type=ctype, sizes=sorted(sizes), intervals=intervals, interval=0.02,
                steps_per_interval=steps, errors_last={m: [float(x) for x in e] for m, e in last.items()},
                velocity_l2_orders=conv.orders(last, 11), pressure_l2_orders=conv.orders(last, 15),
                seconds=time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["fem", "multibody", "ensemble", "conv", "conv-test"])
    ap.add_argument("--type", type=int, default=4)
    ap.add_argument("--sizes", type=str, default="1536,768,512,384,256")
    ap.add_argument("--intervals", type=int, default=50)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--every", type=int, default=1)
    ap.add_argument("--threads", type=int, default=32)
    ap.add_argument("--fp64", action="store_true", help="ensemble: the FP64 solver with previous-solution starts")
    a = ap.parse_args()
    if a.which == "fem":
        r = run_fem()
    elif a.which == "multibody":
        r = run_multibody(steps=a.steps or 200, every=a.every)
    elif a.which == "conv":
        r = run_conv()
    elif a.which == "conv-test":
        r = run_conv_test(a.type, tuple(int(x) for x in a.sizes.split(",")), a.intervals)
    else:
        r = (run_ensemble(steps=a.steps or 10, threads=a.threads, precision=ia.PREC_FP64, guess=ia.GUESS_PREVIOUS)
             if a.fp64 else run_ensemble(steps=a.steps or 10, threads=a.threads))
    r["device"] = ia.device_info()[0]
    print(json.dumps(r))


if __name__ == "__main__":
    main()
