"""Benchmark of the MI355X IncRMT hot path: fluid_2d::step_forward at 1024^2.

Workload: the `ftest simple-spin` rotor physics (ftest.cc:116-143) on a
1024x1024 cell grid, the full per-timestep path (Godunov advection, MAC and
FEM multigrid projections each converged to relative residual <= 1e-8,
level-set extension and reinitialisation), synthetic initial state as the
reference initialises it. One step = one step_forward of adt = (4pi/240)/l.

Prints ONE JSON line. value = cell-steps/s = m*n*K / seconds over K timed
steps (state resident in HBM, no host transfers inside the timed region).
The headline solver is mixed precision (FP32 V-cycles in FP64 refinement)
started from the quadratic extrapolation corrected by the learned guess,
whose networks are first trained (untimed) on the run's own projections;
the other solvers are reported beside it. The roofline object is for the
fine-level smoother of the projections, timed with HIP events on the
library's stream in a further pass of the same K steps; cpu_baseline is the
project's OpenMP CPU oracle on a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def gpu_run(m, steps, warmup, precision, timing=False, guess=0, launch=None, params=None, cost_aware=False):
    import incrmt_amd as ia
    f = ia.simple_spin(m, rel_tol=1e-8, precision=precision, guess=guess)
    if launch is not None:  # otherwise the library's default for the size
        f.set_launch(launch)
    if params is not None:  # trained learned-guess networks (MAC, FEM)
        for k in (ia.MG_MAC, ia.MG_FEM):
            f.set_guess_params(k, params[k])
    if cost_aware:  # the learned start only where it saves more than it costs
        f.set_guess_cost(1)
    l, adt = f.timestep_select(4 * math.pi / 240)
    f.steps(adt, warmup)
    calib = 0
    if cost_aware:
        # untimed, like the networks' training: steps until both projections'
        # cost-aware rule has made its first decision (three learned and one
        # control solve), so that the timed window is the steady state a long
        # run spends its time in, not the first probe
        def decided(k):
            c = f.guess_cost_stats(k)
            return c["off_for_cost"] or (c["learned_counted"] >= 3 and c["control_counted"] >= 1)
        while calib < 16 and not all(decided(k) for k in (ia.MG_MAC, ia.MG_FEM)):
            f.steps(adt, 1)
            calib += 1
    f.sync()
    s0 = f.stats()
    t0 = time.perf_counter()
    f.steps(adt, steps)
    f.sync()
    dt = time.perf_counter() - t0
    s1 = f.stats()
    shapes = f.describe()
    out = dict(seconds=dt, ms_per_step=dt / steps * 1e3, cell_steps_per_s=m * m * steps / dt,
               mac_cycles_per_step=(s1["total_mac_cycles"] - s0["total_mac_cycles"]) / steps,
               fem_cycles_per_step=(s1["total_fem_cycles"] - s0["total_fem_cycles"]) / steps,
               mac_rel_res=s1["mac_rel_res"], fem_rel_res=s1["fem_rel_res"],
               reinit_passes=s1["reinit_passes"], extend_passes=s1["extend_passes"],
               smoother_shapes=shapes)
    if cost_aware:
        out["policy_calibration_steps"] = calib
    # projection-only time: a second pass of the same steps with HIP events
    # around the two projections (their starts included)
    f.time_projections(True)
    f.steps(adt, steps)
    out["projection_ms_per_step"] = f.stats()["proj_ms"] / steps
    f.time_projections(False)
    if guess == ia.GUESS_LEARNED:
        out["learned_start"] = {("mac" if k == ia.MG_MAC else "fem"): dict(f.guess_stats(k), **f.guess_cost_stats(k))
                                for k in (ia.MG_MAC, ia.MG_FEM)}
    if timing:
        # a further pass with HIP events around every fine-level smoother launch
        f.time_smoothers(True)
        f.steps(adt, steps)
        f.sync()
        st = f.stats()
        out["smoother_ms"] = st["smoother_ms"]
        out["smoother_launches"] = st["smoother_launches"]
        out["smoother_bytes"] = st["smoother_bytes"]
        out["mac_cycles_timed"] = st["total_mac_cycles"] - s1["total_mac_cycles"]
        out["fem_cycles_timed"] = st["total_fem_cycles"] - s1["total_fem_cycles"]
        f.time_smoothers(False)
        # the same kernels back to back (no event between launches), for
        # comparison with a profiler's per-dispatch durations
        mac_ms, mac_n = f.time_fine_smoother(ia.MG_MAC, 40, precision)
        fem_ms, fem_n = f.time_fine_smoother(ia.MG_FEM, 40, precision)
        out["back_to_back_us_per_launch"] = {"mac": mac_ms * 1e3, "fem": fem_ms * 1e3}
    f.close()
    return out


def train_guess(m, steps, lr=1e-3):
    """Train the two learned-guess networks online on the solver's own
    projections of the same workload (one Adam step per projection, the
    quadratic start's corrections as targets), untimed for the benchmark."""
    import incrmt_amd as ia
    f = ia.simple_spin(m, rel_tol=1e-8, precision=ia.PREC_MIXED, guess=ia.GUESS_QUADRATIC)
    f.set_guess_training(True, lr, 1e-4)
    l, adt = f.timestep_select(4 * math.pi / 240)
    t0 = time.perf_counter()
    f.steps(adt, steps)
    f.sync()
    dt = time.perf_counter() - t0
    params = [f.guess_params(k) for k in (ia.MG_MAC, ia.MG_FEM)]
    st = {("mac" if k == ia.MG_MAC else "fem"): f.guess_stats(k) for k in (ia.MG_MAC, ia.MG_FEM)}
    f.close()
    return params, dict(steps=steps, projections=2 * steps, seconds=round(dt, 3),
                        ms_per_step_training=round(dt / steps * 1e3, 3),
                        final_loss={k: round(v["loss"], 5) for k, v in st.items()},
                        learning_rate=lr, parameters_per_network=int(params[0].size))


def cpu_info():
    """lscpu's model name and physical core count of this host."""
    import subprocess
    model, cores = "unknown", os.cpu_count() or 1
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
        p = subprocess.run(["lscpu", "-p=Core,Socket"], capture_output=True, text=True, timeout=10).stdout
        pairs = {l for l in p.splitlines() if l and not l.startswith("#")}
        if pairs:
            cores = len(pairs)
    except (OSError, subprocess.SubprocessError):
        pass
    return model, cores


def cpu_baseline(m, steps, warmup=10):
    """The oracle (OpenMP C++ restatement of the reference step, g++ -O3
    -fopenmp) on the host: OMP_NUM_THREADS = the physical cores with
    OMP_PROC_BIND=close and OMP_PLACES=cores, `warmup` untimed steps, then
    `steps` timed; then one thread, one warm-up step and one timed step."""
    model, cores = cpu_info()
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    threads = int(os.environ["OMP_NUM_THREADS"])
    import ctypes
    from tests import oracle_lib as ol
    f = ol.OracleFluid(ol.simple_spin_spec(m), mg_tol=1e-8)
    l, adt = ol.timestep_select(4 * math.pi / 240, f.dt_reg)
    for _ in range(warmup):
        f.step(adt)
    t0 = time.perf_counter()
    for _ in range(steps):
        f.step(adt)
    dt = time.perf_counter() - t0
    out = dict(value=round(m * m * steps / dt, 1), unit="cell-steps/s", cores=threads, kind="port",
               cpu=model, physical_cores=cores,
               sample=f"{steps} oracle steps at {m}x{m} after {warmup} warm-up steps, OpenMP {threads} threads "
                      f"(OMP_PROC_BIND={os.environ['OMP_PROC_BIND']}, OMP_PLACES={os.environ['OMP_PLACES']}), "
                      f"projections to relative residual 1e-8")
    try:
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        f.step(adt)
        t0 = time.perf_counter()
        f.step(adt)
        dt1 = time.perf_counter() - t0
        out["one_thread"] = dict(value=round(m * m / dt1, 1), sample="1 step after 1 warm-up step, 1 thread")
        gomp.omp_set_num_threads(threads)
    except OSError:
        out["one_thread"] = None
    return out


def smoother_probe(m, sweeps):
    """Fine-level smoother sweeps of the two projection systems at the bench
    size (MAC: m x m cells, FEM: (m+1)^2 nodes, variable density of the
    rotor's 3:1 ratio as a smooth blob), timed with HIP events on the
    library stream, plus one device copy of known size (512 MiB read and
    written) for calibrating memory counters. For hardware counter runs."""
    import numpy as np
    import incrmt_amd as ia
    out = {}
    h = 2.0 / m
    x = (np.arange(m) + 0.5) * h - 1
    rho = 1 + 2 * np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / 0.2)
    irho_mac = np.zeros(2 * m * m + 2 * m)
    irho_mac[0:2 * m * m:2] = 1 / rho.ravel()
    irho_mac[1:2 * m * m:2] = 1 / rho.ravel()
    for name, kind, n, irho in (("mac", ia.MG_MAC, m, irho_mac),
                                ("fem", ia.MG_FEM, m + 1, np.ones((m + 1) * (m + 1) + m + 1))):
        s = ia.MgSolver(kind, n, n, h, h)
        s.set_irho(irho)
        for prec, tag in ((ia.PREC_FP64, "fp64"), (ia.PREC_MIXED, "fp32")):
            ms, nl = s.time_smoother(sweeps, prec)
            per_point = (40.0 if kind == ia.MG_MAC else 32.0) * (0.5 if tag == "fp32" else 1.0)
            sweeps_per_launch = sweeps / nl
            gbs = per_point * n * n * sweeps_per_launch / (ms * 1e-3) / 1e9
            out[f"{name}_{tag}"] = dict(us_per_launch=round(ms * 1e3, 2), launches=nl,
                                       sweeps_per_launch=sweeps_per_launch, algorithmic_GBs=round(gbs, 1))
        s.close()
    nbytes = 512 * 1024 * 1024
    ms = ia.copy_probe(nbytes, 4)
    out["copy"] = dict(bytes_each_way=nbytes, ms=round(ms, 4), GBs_read_plus_write=round(2 * nbytes / (ms * 1e-3) / 1e9, 1))
    return out


def measured_traffic(shapes, mac_cycles, fem_cycles, nu=2, fp32=False):
    """HBM bytes per fine-level smoother launch in the timed pass, from the
    calibrated FETCH_SIZE/WRITE_SIZE counts per launch shape measured on the
    same fine levels (profiles/smoother_fine_1024_traffic.json), weighted by
    the launches each system made (pre- and post-smoothing per V-cycle).
    None when a shape in use was not measured."""
    import re
    path = os.path.join(ROOT, "profiles", "smoother_fine_1024_traffic.json")
    if not os.path.exists(path):
        return None
    tab = json.load(open(path))["per_dispatch"]
    total = 0.0
    count = 0
    for kind, part, cycles in ((0, shapes.split(";")[0], mac_cycles), (1, shapes.split(";")[1], fem_cycles)):
        m = re.search(r"(\d+)x(\d+):(\d+)x(\d+)h(\d+)/(\d+)x(\d+)h(\d+)", part)
        g = 6 if fp32 else 3
        sx, sy, h = int(m.group(g)), int(m.group(g + 1)), int(m.group(g + 2))
        name = f"k_smooth_fused<{'float' if fp32 else 'double'}, {kind}, {h}, {sx}, {sy}>"
        if name not in tab:
            return None
        per_launch = h // (2 if kind == 0 else 4)
        n = cycles * 2 * ((nu + per_launch - 1) // per_launch)
        total += n * tab[name]
        count += n
    return round(total / count) if count else None


def sig(v):
    """Floats to 5 significant digits (residuals of 1e-9 stay visible)."""
    return float(f"{v:.5g}") if isinstance(v, float) else v


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--train-steps", type=int, default=500,
                    help="steps of online training of the learned guess before its timed run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--smoother-probe", type=int, default=0, metavar="SWEEPS",
                    help="only run SWEEPS fine-level smoother sweeps of each system (for counter runs)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the mixed+quadratic variant (with --learned: mixed+learned, after the "
                         "untimed training), --warmup then --steps steps (may be 0), no timing passes: "
                         "two profiler runs of it differ by exactly the steps")
    ap.add_argument("--learned", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): the path would shard as row slabs "
              "with halo exchange, which is not built", file=sys.stderr)
        sys.exit(2)
    import incrmt_amd as ia
    m = a.size
    if a.smoother_probe:
        print(json.dumps(smoother_probe(m, a.smoother_probe)))
        return
    if a.headline_only:
        params = train_guess(m, a.train_steps)[0] if a.learned else None
        f = ia.simple_spin(m, rel_tol=1e-8, precision=ia.PREC_MIXED,
                           guess=ia.GUESS_LEARNED if a.learned else ia.GUESS_QUADRATIC)
        if params is not None:
            for k in (ia.MG_MAC, ia.MG_FEM):
                f.set_guess_params(k, params[k])
            f.set_guess_cost(1)
        l, adt = f.timestep_select(4 * math.pi / 240)
        f.steps(adt, a.warmup)
        f.sync()
        t0 = time.perf_counter()
        if a.steps:
            f.steps(adt, a.steps)
        f.sync()
        dt = time.perf_counter() - t0
        print(json.dumps(dict(steps=a.steps, warmup=a.warmup, seconds=round(dt, 6), stats=f.stats())))
        f.close()
        return
    # plain FP64 multigrid warm-started from the previous step's solution as
    # the reference does (mgs_common.hh:48-50), against FP32 V-cycles in FP64
    # refinement started from a prediction in time: the linear extrapolation
    # of the last two solutions (mgs_base::initial_guess with mg_lin_t_interp,
    # mgs_common.cc:33-66), the quadratic one of the last three, and the
    # quadratic start corrected by the learned guess (the headline: the
    # networks are first trained on this workload's own projections, untimed)
    fp64 = gpu_run(m, a.steps, a.warmup, ia.PREC_FP64, timing=True, guess=ia.GUESS_PREVIOUS)
    mixed = gpu_run(m, a.steps, a.warmup, ia.PREC_MIXED, guess=ia.GUESS_PREVIOUS)
    pred = gpu_run(m, a.steps, a.warmup, ia.PREC_MIXED, guess=ia.GUESS_EXTRAPOLATE)
    quad = gpu_run(m, a.steps, a.warmup, ia.PREC_MIXED, guess=ia.GUESS_QUADRATIC)
    # the same with the V-cycles replayed from captured HIP graphs (the
    # library's default below 512^2 unknowns; above it launches them kernel
    # by kernel, as the runs above do)
    graph = gpu_run(m, a.steps, a.warmup, ia.PREC_MIXED, guess=ia.GUESS_QUADRATIC, launch=ia.LAUNCH_GRAPH)
    params, training = train_guess(m, a.train_steps)
    # the learned start used on every solve, and (the headline) only where
    # it saves more V-cycles than its network costs (measured in V-cycles)
    always = gpu_run(m, a.steps, a.warmup, ia.PREC_MIXED, guess=ia.GUESS_LEARNED, params=params)
    learned = gpu_run(m, a.steps, a.warmup, ia.PREC_MIXED, timing=True, guess=ia.GUESS_LEARNED, params=params,
                      cost_aware=True)
    best = learned
    best_name = "mixed+learned_guess (cost-aware)"
    # roofline of the fine-level smoother: algorithmic bytes per sweep are,
    # per point, z, b and the coefficients read and z written: MAC
    # red-black 8+8+16+8 = 40 B/cell in FP64 (two face inverse densities),
    # 20 in FP32; FEM 4-colour 8+8+8+8 = 32 B/node (one cell inverse
    # density), 16 in FP32; the library sums them over the timed sweeps and
    # times each launch with HIP events. The headline's V-cycles smooth in
    # FP32 (roofline), the FP64 baseline's in FP64 (roofline_fp64).
    def smoother_roofline(r, fp32):
        launches = r["smoother_launches"]
        achieved = r["smoother_bytes"] / (r["smoother_ms"] * 1e-3) / 1e9
        t = "float" if fp32 else "double"
        return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=measured_traffic(r["smoother_shapes"], r["mac_cycles_timed"], r["fem_cycles_timed"],
                                             fp32=fp32),
                    kernel=f"fine-level projection smoother k_smooth_fused<{t}> (MAC and FEM)",
                    avg_launch_us=round(r["smoother_ms"] / launches * 1e3, 2),
                    back_to_back_us_per_launch={k: round(v, 2) for k, v in r["back_to_back_us_per_launch"].items()},
                    algorithmic_bytes_per_launch=round(r["smoother_bytes"] / launches))
    roofline = smoother_roofline(learned, True)
    roofline_fp64 = smoother_roofline(fp64, False)
    name, hbm = ia.device_info()
    skip = ("back_to_back_us_per_launch", "learned_start")
    line = {
        "metric": "cell-steps/s at 1024^2 (plain FP64 MG vs mixed+learned), rel. div <=1e-8",
        "value": round(best["cell_steps_per_s"], 1),
        "unit": "cell-steps/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(best["ms_per_step"], 3),
        "projection_ms_per_step": round(best["projection_ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32+bf16",
        "data": "synthetic (simple-spin rotor initial state); learned-guess networks trained on the run's own projections",
        "config": {"workload": f"simple-spin rotor, full step_forward, {m}x{m} cells, projections to rel. residual 1e-8",
                   "solver": best_name, "baseline_solver": "fp64 multigrid, previous-solution warm start",
                   "grid": [m, m]},
        "variants": {"fp64": {k: sig(v) for k, v in fp64.items() if k not in skip},
                     "mixed": {k: sig(v) for k, v in mixed.items()},
                     "mixed+predicted_guess": {k: sig(v) for k, v in pred.items()},
                     "mixed+quadratic_guess": {k: sig(v) for k, v in quad.items()},
                     "mixed+quadratic_guess, graph replay": {k: sig(v) for k, v in graph.items()
                                                             if k != "smoother_shapes"},
                     "mixed+learned_guess, every solve": {k: sig(v) for k, v in always.items() if k not in skip},
                     "mixed+learned_guess (cost-aware)": {k: sig(v) for k, v in learned.items() if k not in skip}},
        "learned_guess": {"training": training,
                          "learned_start": {k: {kk: sig(vv) for kk, vv in v.items()}
                                            for k, v in learned["learned_start"].items()},
                          "learned_start_every_solve": {k: {kk: sig(vv) for kk, vv in v.items()}
                                                        for k, v in always["learned_start"].items()}},
        "speedup_vs_plain_fp64": round(best["cell_steps_per_s"] / fp64["cell_steps_per_s"], 3),
        "speedup_quadratic_vs_plain_fp64": round(quad["cell_steps_per_s"] / fp64["cell_steps_per_s"], 3),
        "roofline": roofline,
        "roofline_fp64": roofline_fp64,
        "device": name,
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(m, a.cpu_steps)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
