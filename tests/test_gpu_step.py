"""GPU fluid_2d::step_forward vs the oracle, per step.

Both sides solve the projections to a relative residual far below the
stated tolerances (GPU 1e-12, oracle 1e-13). Stated per-step tolerances:
u, v, p relative L2 <= 1e-6; X, Y max |difference| <= 1e-9; band status
codes equal on at least 99.5% of the band points and phi within 0.05 dx
(the GPU band rebuild equals the reference's heap march except on exact
phi ties, which the mirror-symmetric initial rotor has). After the first
steps the differences are roundoff plus the solver tolerance (~1e-10).
"""
import math

import numpy as np
import pytest

import incrmt_amd as ia
from tests import divergence as dv
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


def compare(g, o, m, dx, tag):
    inner = (slice(2, -2), slice(2, -2))
    for nm in ("u", "v", "p"):
        r = rel(g.get(nm)[inner], o.get(nm)[inner])
        assert r <= 1e-6, f"{tag}: {nm} relative L2 {r:.3e}"
    for k in range(g.nobj):
        for nm in ("X", "Y"):
            d = np.abs(g.get(nm, k) - o.get(nm, k))[inner].max()
            assert d <= 1e-9, f"{tag}: {nm}[{k}] max diff {d:.3e}"
        sg, so = g.get("s", k), o.get("s", k)
        band = (so >= 2) & (so <= 5)
        if not band.any():  # no interface (obj_full): statuses and phi agree everywhere
            assert np.array_equal(sg, so), f"{tag}: status mismatch"
            assert np.abs(g.get("phi", k) - o.get("phi", k)).max() <= 0.05 * dx, f"{tag}: phi diff"
            continue
        assert (sg[band] != so[band]).mean() <= 0.005, f"{tag}: status mismatch"
        dphi = np.abs(g.get("phi", k) - o.get("phi", k))[band & (sg == so)].max()
        assert dphi <= 0.05 * dx, f"{tag}: phi diff {dphi:.3e}"


def gpu_from_spec(spec, tol=1e-12, precision=ia.PREC_FP64, guess=ia.GUESS_PREVIOUS):
    f = ia.Fluid2D(spec["m"], spec["n"], visc=spec["visc"], rhof=spec["rhof"], tw=spec["tw"], pinw=spec["pinw"],
                   ax=spec["ax"], bx=spec["bx"], ay=spec["ay"], by=spec["by"])
    for kind, p, G, rhos, evm in spec["objects"]:
        if kind == ol.SPIN:
            f.add_spin(*p, G=G, rhos=rhos, ev_trans_mult=evm)
        else:
            f.add_circle(p[0], p[1], p[2], p[3], G=G, rhos=rhos, ev_trans_mult=evm)
    f.set_solver(rel_tol=tol, max_cycles=200, precision=precision)
    f.set_guess(guess)
    f.initialize(spec["ev_mult"], spec["dt_pad"], spec["dt_ev_pad"])
    return f


# (precision, starting iterate): the reference's FP64 warm start, and the
# benchmark's solver (FP32 V-cycles in FP64 refinement from the quadratic
# extrapolation in time)
SOLVERS = [(ia.PREC_FP64, ia.GUESS_PREVIOUS), (ia.PREC_MIXED, ia.GUESS_QUADRATIC)]


@pytest.mark.parametrize("solver", SOLVERS, ids=["fp64", "mixed+quadratic"])
@pytest.mark.parametrize("m", [64, 128])
def test_simple_spin_per_step(m, solver):
    spec = ol.simple_spin_spec(m)
    g, o = gpu_from_spec(spec, precision=solver[0], guess=solver[1]), ol.OracleFluid(spec, mg_tol=1e-13)
    assert g.dt_reg == o.dt_reg
    l, adt = g.timestep_select(4 * math.pi / 240)
    dx = 2.0 / m
    compare(g, o, m, dx, "init")
    for k in range(1, 21):
        g.step_forward(adt)
        o.step(adt)
        compare(g, o, m, dx, f"step {k}")
    assert g.time == pytest.approx(o.time, rel=1e-15)


def circles_spec(m, centres):
    # sediment-style solids: rho_s/rho_f = 2, G = 1, mu = 0.01, gravity 1
    return dict(m=m, n=m, ax=-1.0, bx=1.0, ay=-1.0, by=1.0, visc=0.01, rhof=1.0, tw=2.5, pinw=2.5,
                w_scale=3.0, w_force=20.0, ev_mult=0.8, dt_pad=0.4, dt_ev_pad=0.8,
                objects=[(ol.CIRCLE, [cx, cy, r, 1.0, 0, 0], 1.0, 2.0, 1.0) for cx, cy, r in centres])


@pytest.mark.parametrize("centres", [[(0.13, 0.07, 0.3)], [(-0.31, 0.02, 0.27), (0.29, -0.04, 0.3)]])
def test_sedimenting_circles_per_step(centres):
    m = 64
    spec = circles_spec(m, centres)
    g, o = gpu_from_spec(spec), ol.OracleFluid(spec, mg_tol=1e-13)
    dx = 2.0 / m
    for k in range(1, 11):
        g.step_forward(g.dt_reg)
        o.step(o.dt_reg)
        compare(g, o, m, dx, f"step {k}")


def test_projected_divergence_matches_oracle():
    m = 64
    spec = ol.simple_spin_spec(m)
    g, o = gpu_from_spec(spec), ol.OracleFluid(spec, mg_tol=1e-13)
    l, adt = g.timestep_select(4 * math.pi / 240)
    for _ in range(5):
        g.step_forward(adt)
        o.step(adt)
    h = 2.0 / m
    dg = dv.mac_div(g.get("ul"), g.get("vd"), m, m, h, h)
    assert np.abs(dg).max() <= 1e-9 * np.abs(g.get("ul")).max() / h
    fg = dv.fem_div(g.get("u"), g.get("v"), m, m, h, h)
    fo = dv.fem_div(o.get("u"), o.get("v"), m, m, h, h)
    assert rel(fg, fo) <= 1e-6


def test_rest_state():
    g = ia.simple_spin(128, omega=0.0, rel_tol=1e-12)
    g.steps(g.dt_reg, 4)
    assert np.abs(g.get("u")).max() < 1e-8 and np.abs(g.get("v")).max() < 1e-8


def test_1024_properties():
    # the bench workload at full size, checked by size-independent properties
    g = ia.simple_spin(1024, rel_tol=1e-8)
    l, adt = g.timestep_select(4 * math.pi / 240)
    g.steps(adt, 3)
    st = g.stats()
    assert st["mac_rel_res"] <= 1e-8 and st["fem_rel_res"] <= 1e-8
    u, v = g.get("u"), g.get("v")
    assert np.isfinite(u).all() and np.isfinite(v).all()
    h = 2.0 / 1024
    dg = dv.mac_div(g.get("ul"), g.get("vd"), 1024, 1024, h, h)
    assert np.abs(dg).max() <= 1e-6 * np.abs(g.get("ul")).max() / h
    # interface layers sit on sign changes: every status-3/4 point has a
    # 4-neighbour of the other sign
    s, phi = g.get("s"), g.get("phi")
    iface = (s == 3) | (s == 4)
    pos = phi > 0
    other = np.zeros_like(iface)
    other[1:, :] |= pos[1:, :] != pos[:-1, :]
    other[:-1, :] |= pos[:-1, :] != pos[1:, :]
    other[:, 1:] |= pos[:, 1:] != pos[:, :-1]
    other[:, :-1] |= pos[:, :-1] != pos[:, 1:]
    assert np.all(other[iface])
    # the anchor turns the rotor counter-clockwise (theta = pi(1-cos t),
    # object.cc:175-181): the solid's angular momentum about the pivot is positive
    x = -1 + h * (np.arange(1028) - 1.5)
    X, Y = np.meshgrid(x, x)
    solid = phi < 0
    assert (X * v - Y * u)[solid].sum() > 0


def test_solver_failure_is_reported():
    g = ia.simple_spin(64)
    g.set_solver(rel_tol=1e-15, max_cycles=1)
    with pytest.raises(ia.IncrmtError) as e:
        g.step_forward(g.dt_reg)
    assert e.value.code == ia.ERR_NOT_CONVERGED


@pytest.mark.parametrize("precision", [ia.PREC_FP64, ia.PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("guess", [ia.GUESS_EXTRAPOLATE, ia.GUESS_QUADRATIC])
def test_extrapolated_guess_same_solution(guess, precision):
    # starting the projections from the linear (mg_lin_t_interp) or
    # quadratic extrapolation in time changes the iterates, not the answer:
    # per-step parity with the oracle as above, fewer cycles
    m = 64
    spec = ol.simple_spin_spec(m)
    g, o = gpu_from_spec(spec, precision=precision), ol.OracleFluid(spec, mg_tol=1e-13)
    g.set_guess(guess)
    base = gpu_from_spec(spec, precision=precision)
    l, adt = g.timestep_select(4 * math.pi / 240)
    dx = 2.0 / m
    for k in range(1, 16):
        g.step_forward(adt)
        o.step(adt)
        base.step_forward(adt)
        compare(g, o, m, dx, f"step {k}")
    sg, sb = g.stats(), base.stats()
    assert sg["total_mac_cycles"] + sg["total_fem_cycles"] < sb["total_mac_cycles"] + sb["total_fem_cycles"]


@pytest.mark.parametrize("m", [64, 128])
def test_benchmark_solver_at_its_tolerance(m):
    # the benchmark's exact configuration (mixed precision, quadratic start,
    # projections stopped at relative residual 1e-8) against the oracle
    # solved to 1e-13, 20 steps. Stated tolerances for a 1e-8 solve: u, v
    # relative L2 <= 1e-6, p relative L2 <= 1e-5, X, Y max |difference|
    # <= 1e-9; the band as above.
    spec = ol.simple_spin_spec(m)
    g = gpu_from_spec(spec, tol=1e-8, precision=ia.PREC_MIXED, guess=ia.GUESS_QUADRATIC)
    o = ol.OracleFluid(spec, mg_tol=1e-13)
    l, adt = g.timestep_select(4 * math.pi / 240)
    dx = 2.0 / m
    inner = (slice(2, -2), slice(2, -2))
    worst = {}
    for k in range(1, 21):
        g.step_forward(adt)
        o.step(adt)
        st = g.stats()
        assert st["mac_rel_res"] <= 1e-8 and st["fem_rel_res"] <= 1e-8
        for nm, tol in (("u", 1e-6), ("v", 1e-6), ("p", 1e-5)):
            r = rel(g.get(nm)[inner], o.get(nm)[inner])
            worst[nm] = max(worst.get(nm, 0), r)
            assert r <= tol, f"step {k}: {nm} relative L2 {r:.3e}"
        for nm in ("X", "Y"):
            d = np.abs(g.get(nm) - o.get(nm))[inner].max()
            assert d <= 1e-9, f"step {k}: {nm} max diff {d:.3e}"
    print(f"benchmark solver at {m}^2, worst relative L2 over 20 steps:", worst)


def test_quadratic_start_long_run():
    # 1200 steps at 64^2 from the quadratic start: the iterates' constant
    # part stays bounded (the extrapolated start is kept mean-free) and the
    # fields stay with those of the previous-solution start
    spec = ol.simple_spin_spec(64)
    q = gpu_from_spec(spec, tol=1e-10, precision=ia.PREC_MIXED, guess=ia.GUESS_QUADRATIC)
    b = gpu_from_spec(spec, tol=1e-10, precision=ia.PREC_FP64, guess=ia.GUESS_PREVIOUS)
    l, adt = q.timestep_select(4 * math.pi / 240)
    inner = (slice(2, -2), slice(2, -2))
    for k in range(12):
        q.steps(adt, 100)
        b.steps(adt, 100)
        for nm in ("u", "v", "p"):
            r = rel(q.get(nm)[inner], b.get(nm)[inner])
            assert r <= 1e-6, f"after {100 * (k + 1)} steps: {nm} relative L2 {r:.3e}"
    # switching the mode clears the history: a linear start after the
    # quadratic run is as good as a fresh one
    q.set_guess(ia.GUESS_EXTRAPOLATE)
    q.steps(adt, 5)
    b.steps(adt, 5)
    assert rel(q.get("p")[inner], b.get("p")[inner]) <= 1e-6
    assert q.stats()["mac_cycles"] <= b.stats()["mac_cycles"] + 2


def test_launch_mode_before_initialize_is_kept():
    # LAUNCH_DIRECT asked for before initialize survives the solvers'
    # set-up (which would pick graphs at 64^2); results are bit-identical
    spec = ol.simple_spin_spec(64)
    f = ia.Fluid2D(64, 64, visc=spec["visc"], rhof=spec["rhof"], tw=spec["tw"], pinw=spec["pinw"])
    f.add_spin(*spec["objects"][0][1], G=48.0, rhos=3.0, ev_trans_mult=1.0)
    f.set_launch(ia.LAUNCH_DIRECT)
    f.set_solver(rel_tol=1e-12, max_cycles=200)
    f.initialize(spec["ev_mult"], spec["dt_pad"], spec["dt_ev_pad"])
    g = gpu_from_spec(spec)
    for h in (f, g):
        h.steps(h.dt_reg, 3)
    assert np.array_equal(f.get("p"), g.get("p"))
    assert f.describe().endswith("launch MAC direct, FEM direct")
    assert g.describe().endswith("launch MAC graph, FEM graph")
