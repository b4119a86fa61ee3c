"""Time every fused-smoother launch shape on the fine level of both
projection systems at a ladder of sizes, and print the winners as the
tuned-shape table the multigrid compiles in (kTuned in csrc/mg.hip).

    python -m incrmt_amd.tune_shapes [--sizes 1024 512 256 128 64] [--json out.json] [--calibrate]

Each level is an m x m MAC grid (cells) and an (m+1) x (m+1) FEM grid
(nodes) with the rotor's 3:1 density ratio as a smooth blob. A shape's
time is the median over repeats of the mean launch time of `sweeps`
sweeps (HIP events on the library stream); the winner is the shape with
the smallest time per sweep. Run on the GPU the table is for; it needs
nothing but the built library. --calibrate adds one device copy of
512 MiB read and written, for scaling memory counters of a profiler run
of the same command.
"""
import argparse
import json

import numpy as np

import incrmt_amd as ia

SHAPES = [(64, 16, 4), (64, 8, 4), (64, 4, 4), (64, 16, 2), (64, 8, 2), (64, 4, 2), (64, 8, 8), (64, 4, 8),
          (64, 32, 4), (128, 16, 4), (64, 32, 8), (128, 16, 8), (128, 8, 8), (64, 64, 8), (128, 32, 4),
          (128, 32, 8), (128, 64, 8)]


def blob_irho(m):
    h = 2.0 / m
    x = (np.arange(m) + 0.5) * h - 1
    rho = 1 + 2 * np.exp(-(x[None, :] ** 2 + x[:, None] ** 2) / 0.2)
    irho = np.zeros(2 * m * m + 2 * m)
    irho[0:2 * m * m:2] = 1 / rho.ravel()
    irho[1:2 * m * m:2] = 1 / rho.ravel()
    return irho


def sweep_level(m, sweeps=8, repeats=5, log=print):
    """{(kind, precision): [(shape, us per launch, sweeps per launch), ...]}"""
    h = 2.0 / m
    out = {}
    for kind, n, irho in ((ia.MG_MAC, m, blob_irho(m)), (ia.MG_FEM, m + 1, np.ones((m + 1) * (m + 1) + m + 1))):
        s = ia.MgSolver(kind, n, n, h, h)
        s.set_irho(irho)
        for prec in (ia.PREC_FP64, ia.PREC_MIXED):
            rows = []
            for k in range(len(SHAPES)):
                try:
                    s.time_smoother(sweeps, prec, k)  # warm-up (and validity)
                except ia.IncrmtError:
                    continue
                t = [s.time_smoother(sweeps, prec, k) for _ in range(repeats)]
                us = float(np.median([ms for ms, _ in t])) * 1e3
                rows.append((k, us, sweeps / t[0][1]))
                log(f"{'MAC' if kind == ia.MG_MAC else 'FEM'} {n}^2 {'fp64' if prec == ia.PREC_FP64 else 'fp32'} "
                    f"shape {k:2d} {SHAPES[k][0]}x{SHAPES[k][1]}h{SHAPES[k][2]}: {us:8.2f} us/launch "
                    f"({sweeps / t[0][1]:.0f} sweeps)")
            out[(kind, prec)] = rows
        s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", type=int, nargs="+", default=[1024, 512, 256, 128, 64])
    ap.add_argument("--json", default=None)
    ap.add_argument("--calibrate", action="store_true")
    a = ap.parse_args()
    table, raw = [], {}
    for m in a.sizes:
        res = sweep_level(m, sweeps=8 if m >= 256 else 16)
        for (kind, prec), rows in res.items():
            best = min(rows, key=lambda r: r[1] / r[2])
            n = m if kind == ia.MG_MAC else m + 1
            table.append((kind, prec, n, best[0]))
            raw[f"{'mac' if kind == ia.MG_MAC else 'fem'}_{n}_{'fp64' if prec == ia.PREC_FP64 else 'fp32'}"] = rows
    print("// kind, FP32 sweeps, nx, ny, shape")
    for kind, prec, n, sh in table:
        print(f"    {{{'kMAC' if kind == ia.MG_MAC else 'kFEM'}, {int(prec == ia.PREC_MIXED)}, {n}, {n}, {sh}}},"
              f"  // {SHAPES[sh][0]}x{SHAPES[sh][1]}h{SHAPES[sh][2]}")
    if a.calibrate:
        nbytes = 512 * 1024 * 1024
        ms = ia.copy_probe(nbytes, 4)
        print(f"copy of {nbytes} bytes: {ms:.4f} ms, {2 * nbytes / (ms * 1e-3) / 1e9:.1f} GB/s read + write")
    if a.json:
        json.dump(raw, open(a.json, "w"), indent=1)


if __name__ == "__main__":
    main()
