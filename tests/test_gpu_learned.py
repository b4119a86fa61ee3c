"""The learned projection guess (INCRMT_GUESS_LEARNED, the hook of
mgs_base::initial_guess, mgs_common.cc:33-66) on the GPU.

- An untrained network's start still gives the oracle's step to the stated
  per-step tolerances (the start only moves the iterate; every projection
  stops at the same relative residual).
- A fallback forced from the host (tolerance 0) takes every learned start
  back to the quadratic start with FP64 V-cycles, the results stay with the
  oracle, and the breaker switches the learned start off after 8 misses.
- Trained online on the solver's own projections (the quadratic start's
  corrections), the learned start needs fewer V-cycles than the quadratic
  start at the same relative residual (1e-8)."""
import math

import numpy as np
import pytest

import incrmt_amd as ia
from tests import oracle_lib as ol
from tests.test_gpu_step import compare, gpu_from_spec

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fallback_tol", [1e-4, 0.0], ids=["default", "forced-fallback"])
def test_learned_start_matches_oracle(fallback_tol):
    m = 64
    spec = ol.simple_spin_spec(m)
    g = gpu_from_spec(spec, tol=1e-12, precision=ia.PREC_MIXED, guess=ia.GUESS_LEARNED)
    g.set_guess_training(False, 1e-3, fallback_tol)
    o = ol.OracleFluid(spec, mg_tol=1e-13)
    l, adt = g.timestep_select(4 * math.pi / 240)
    for k in range(1, 16):
        g.step_forward(adt)
        o.step(adt)
        compare(g, o, m, 2.0 / m, f"step {k}")
    for kind in (ia.MG_MAC, ia.MG_FEM):
        st = g.guess_stats(kind)
        assert st["learned_solves"] >= 8          # from the third solve on
        if fallback_tol == 0.0:
            assert st["fallbacks"] == st["learned_solves"] == 8 and st["off"]


def run(m, steps, guess, params=None, train=False):
    f = ia.simple_spin(m, rel_tol=1e-8, precision=ia.PREC_MIXED, guess=guess)
    if train:
        f.set_guess_training(True, 1e-3, 1e-4)
    if params is not None:
        for k in (ia.MG_MAC, ia.MG_FEM):
            f.set_guess_params(k, params[k])
    l, adt = f.timestep_select(4 * math.pi / 240)
    cyc = np.zeros(2)
    for _ in range(steps):
        f.step_forward(adt)
        st = f.stats()
        assert st["mac_rel_res"] <= 1e-8 and st["fem_rel_res"] <= 1e-8   # equal residual
        cyc += (st["mac_cycles"], st["fem_cycles"])
    out = dict(cycles=cyc, stats=[f.guess_stats(k) for k in (ia.MG_MAC, ia.MG_FEM)],
               params=[f.guess_params(k) for k in (ia.MG_MAC, ia.MG_FEM)] if (train or params is not None) else None)
    f.close()
    return out


def test_trained_guess_saves_v_cycles():
    m = 128
    tr = run(m, 1200, ia.GUESS_QUADRATIC, train=True)
    assert tr["stats"][1]["train_steps"] >= 1190
    quad = run(m, 300, ia.GUESS_QUADRATIC)
    learned = run(m, 300, ia.GUESS_LEARNED, params=tr["params"])
    print(f"V-cycles over 300 steps at {m}^2 (MAC, FEM): quadratic {quad['cycles']}, learned {learned['cycles']}; "
          f"training loss {tr['stats'][0]['loss']:.4f} / {tr['stats'][1]['loss']:.4f}; "
          f"learned-start stats {learned['stats']}")
    assert learned["cycles"][1] <= 0.95 * quad["cycles"][1]       # FEM
    assert learned["cycles"].sum() < quad["cycles"].sum()


def test_cost_aware_use():
    """With the learned start's cost given as 100 V-cycles (mode 2), which no
    start can save, the policy runs one control solve (quadratic start) in
    four and switches the learned start off once three learned and one
    control solve are counted: 3 learned solves per projection, then none.
    With the cost measured (mode 1) it is a positive number of V-cycles.
    Every projection still reaches its residual."""
    f = ia.simple_spin(64, rel_tol=1e-8, precision=ia.PREC_MIXED, guess=ia.GUESS_LEARNED)
    f.set_guess_training(False, 1e-3, 1.0)   # every learned start accepted: no misses
    f.set_guess_cost(2, 100.0)
    l, adt = f.timestep_select(4 * math.pi / 240)
    for _ in range(30):
        f.step_forward(adt)
        st = f.stats()
        assert st["mac_rel_res"] <= 1e-8 and st["fem_rel_res"] <= 1e-8
    for kind in (ia.MG_MAC, ia.MG_FEM):
        cs, gs = f.guess_cost_stats(kind), f.guess_stats(kind)
        assert cs["off_for_cost"] and gs["learned_solves"] == 3, (cs, gs)
    f.set_guess_cost(1)
    for _ in range(2):
        f.step_forward(adt)
    for kind in (ia.MG_MAC, ia.MG_FEM):
        cs = f.guess_cost_stats(kind)
        assert cs["cost"] > 0 and not cs["off_for_cost"], cs
    f.close()
