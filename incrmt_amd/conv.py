"""Resolution comparison of the conv_test driver (conv_test.cc:14-194) for
boxes with walls or periodic directions: fluid_2d::l_comparison
(fluid_2d_io.cc:186-281)
restated with array operations, and a ladder of runs of one case at
several resolutions, each compared with the finest, from which the
convergence order follows.

l_comparison(coarse, fine) compares a coarse run with a finer one whose
grid is an integer multiple: the velocity at coarse cell centres against
the fine velocity interpolated there (the centre value for odd ratios, the
4-point cubic stencil -1/16, 9/16, 9/16, -1/16 for even ones), and the
pressure at coarse grid corners against the fine pressure at the same
corners (boundary corners weighted 1/2, corners of the box 1/4). Points are
split by the material there (fluid: min phi > eps, solid: < -eps,
transition otherwise). The 32 numbers are, in groups of four (fluid,
solid, transition, total): velocity point counts, pressure point counts,
velocity L2, pressure L2, velocity L1, pressure L1 (all divided by m n;
L2 as square roots), velocity L-infinity, pressure maximum (the signed
maximum, as the reference takes it).
"""
import math

import numpy as np

_STE = np.array([1.0, -1.0 / 16.0, 9.0 / 16.0, 9.0 / 16.0, -1.0 / 16.0])


def _state(f):
    """The fields l_comparison reads, from an incrmt_amd.Fluid2D (or any
    object with the same get/m/n/nobj/eps members, and x_prd/y_prd when
    periodic)."""
    phis = [f.get("phi", k) for k in range(f.nobj)]
    return dict(m=f.m, n=f.n, u=f.get("u"), v=f.get("v"), p=f.get("p"), phi=phis, eps=f.eps,
                px=bool(getattr(f, "x_prd", False)), py=bool(getattr(f, "y_prd", False)))


def _min_phi(st):
    if not st["phi"]:
        return np.full_like(st["u"], np.finfo(np.float64).max)
    return np.minimum.reduce(st["phi"])


def l_comparison(coarse, fine):
    """fluid_2d::l_comparison of `coarse` against `fine` (dicts from _state).
    Returns the 32 error numbers."""
    m, n = coarse["m"], coarse["n"]
    pm, pn = fine["m"], fine["n"]
    if pm % m or pn % n:
        raise ValueError("grid dimension is not a multiple of the finer grid")
    dm, dn = pm // m, pn // n
    if dm & 1:
        li, ui, stx = dm // 2, dm // 2 + 1, _STE[0:1]
    else:
        li, ui, stx = dm // 2 - 2, dm // 2 + 2, _STE[1:5]
    if dn & 1:
        lj, uj, sty = dn // 2, dn // 2 + 1, _STE[0:1]
    else:
        lj, uj, sty = dn // 2 - 2, dn // 2 + 2, _STE[1:5]
    eps = coarse["eps"]
    o = np.zeros(24)
    # velocity at the coarse cell centres (i < m, j < n)
    fu, fv = fine["u"], fine["v"]
    iu = np.zeros((n, m))
    iv = np.zeros((n, m))
    for a, jj in enumerate(range(lj, uj)):
        for b, ii in enumerate(range(li, ui)):
            w = sty[a] * stx[b]
            rows = slice(2 + jj, 2 + jj + dn * n, dn)
            cols = slice(2 + ii, 2 + ii + dm * m, dm)
            iu += w * fu[rows, cols]
            iv += w * fv[rows, cols]
    du = iu - coarse["u"][2:n + 2, 2:m + 2]
    dv = iv - coarse["v"][2:n + 2, 2:m + 2]
    mp = _min_phi(coarse)
    cls_c = np.where(mp[2:n + 2, 2:m + 2] > eps, 0, np.where(mp[2:n + 2, 2:m + 2] < -eps, 1, 2))
    delsq = du * du + dv * dv
    dl = np.sqrt(delsq)
    for c in range(3):
        sel = cls_c == c
        o[c] += sel.sum()
        o[6 + c] += delsq[sel].sum()
        o[12 + c] += dl[sel].sum()
        if sel.any():
            o[18 + c] = max(o[18 + c], dl[sel].max())
    # pressure at the coarse corners (i <= m, j <= n; i < m, j < n along a
    # periodic direction, whose corners all weigh 1)
    px, py = coarse.get("px", False), coarse.get("py", False)
    J, I = np.mgrid[0:(n if py else n + 1), 0:(m if px else m + 1)]
    tf = np.where(py | ((J > 0) & (J < n)), 1.0, 0.5)
    tf2 = np.where(px | ((I > 0) & (I < m)), tf, 0.5 * tf)
    ml = m + 4
    # the stencil of the top row and right column of corners reaches one row
    # and column past the ghosted grid (the reference reads past its level-set
    # array there); those points count as far from every solid
    mpf = np.concatenate([mp.ravel(), np.full(3 * ml, np.finfo(np.float64).max)])
    mph = np.zeros(J.shape)
    base = ml * (2 + J) + (I + 2)
    with np.errstate(over="ignore", invalid="ignore"):
        for jj in range(-1, 3):
            for ii in range(-1, 3):
                mph += _STE[jj + 2] * _STE[ii + 2] * mpf[base + jj * ml + ii]
    cls_p = np.where(mph > eps, 0, np.where(mph < -eps, 1, 2))
    delp = fine["p"][2 + dn * J, 2 + dm * I] - coarse["p"][2 + J, 2 + I]
    for c in range(3):
        sel = cls_p == c
        o[3 + c] += sel.sum()
        o[9 + c] += (delp[sel] ** 2 * tf2[sel]).sum()
        o[15 + c] += (np.abs(delp[sel]) * tf2[sel]).sum()
        if sel.any():
            o[21 + c] = max(o[21 + c], delp[sel].max())
    err = np.zeros(32)
    fac = 1.0 / (m * n)
    for a in range(6):
        for c in range(3):
            err[4 * a + c] = o[3 * a + c] * fac
        err[4 * a + 3] = err[4 * a] + err[4 * a + 1] + err[4 * a + 2]
    for a in range(6, 8):
        for c in range(3):
            err[4 * a + c] = o[3 * a + c]
        err[4 * a + 3] = max(err[4 * a], err[4 * a + 1], err[4 * a + 2])
    err[8:16] = np.sqrt(err[8:16])
    return err


def ladder(make, sizes, duration):
    """Run make(m) (an initialised Fluid2D) for `duration` at each size with
    the largest timestep below its dt_reg that divides the duration
    (timestep_select), then compare every size with the finest. Returns
    {m: 32 errors} for all sizes but the finest, and the steps taken."""
    states = {}
    steps = {}
    for m in sorted(sizes):
        f = make(m)
        l, adt = f.timestep_select(duration)
        f.steps(adt, l)
        states[m] = _state(f)
        steps[m] = l
        f.close()
    finest = states[max(sizes)]
    return {m: l_comparison(states[m], finest) for m in sorted(sizes) if m != max(sizes)}, steps


def orders(errs, index=11):
    """Observed orders log2(e(h)/e(h/2)) of one error quantity (default: the
    total velocity L2, err[11]) between successive sizes."""
    ms = sorted(errs)
    return [math.log(errs[a][index] / errs[b][index]) / math.log(b / a) for a, b in zip(ms[:-1], ms[1:])]


# conv_test's test types (conv_test.cc:59-88): solid density and fluid
# viscosity per type 0..7; types 0 and 1 are fluid only, 3 carries
# obj_square(-0.1,0,0.6) and 4..7 obj_circle(-0.1,0,0.6), with G = 1.
# Type 2 (obj_full with periodic solid boundaries) is not built here.
CONV_RHOS = (1, 1, 1, 1, 3, 1, 1, 1)
CONV_VISC = (1e-3, 4e-3, 1e-20, 1e-3, 1e-3, 1e-3, 4e-3, 16e-3)


def conv_test_case(m, ctype, coarsest=280, rel_tol=1e-10, precision=0):
    """One of conv_test's simulations (conv_test.cc:98-128) on the GPU: the
    periodic [-1,1]^2 box with sim_velocity_pulses, initialized with
    ev_mult 0.4 and the timestep paddings 0.4 and 0.8 scaled by
    coarsest / m (the reference scales by its coarsest grid, 280), so the
    timestep falls as h^2. Its solids leave the pulses field inside them
    (mat_const(G, rhos, ev_trans_mult, false), conv_test.cc:108)."""
    import incrmt_amd as ia
    if ctype not in range(8):
        raise ValueError("conv_test types are 0..7")
    f = ia.Fluid2D(m, m, visc=CONV_VISC[ctype], rhof=1.0, tw=2.5, pinw=2.5, x_prd=True, y_prd=True)
    if ctype == 2:  # obj_full with solid periodic boundaries (conv_test.cc:83-85, 120)
        f.add_full(G=1.0, rhos=float(CONV_RHOS[ctype]), ev_trans_mult=1.0)
        f.set_solid_periodic(True)
    elif ctype == 3:
        f.add_square(-0.1, 0.0, 0.6, 0.0, G=1.0, rhos=float(CONV_RHOS[ctype]), ev_trans_mult=1.0)
    elif ctype >= 4:
        f.add_circle(-0.1, 0.0, 0.6, 0.0, G=1.0, rhos=float(CONV_RHOS[ctype]), ev_trans_mult=1.0)
    if ctype >= 2:
        f.set_solid_velocity(0, False)
    f.set_sim_velocity(ia.SIM_VELOCITY_PULSES)
    f.set_solver(rel_tol=rel_tol, precision=precision)
    f.initialize(0.4, 0.4 * coarsest / m, 0.8 * coarsest / m)
    return f


def conv_test_run(ctype, sizes, intervals=50, interval=0.02, coarsest=280, rel_tol=1e-10, callback=None):
    """conv_test's main loop (conv_test.cc:130-175): every size advances
    interval by interval (timestep_select(interval) steps each), and at
    time 0 and after each interval every size but the finest is compared
    with the finest by l_comparison. Returns ([{m: 32 errors}] for
    intervals 0..intervals, {m: steps per interval})."""
    sims = {m: conv_test_case(m, ctype, coarsest, rel_tol) for m in sorted(sizes)}
    finest = max(sizes)
    sel = {m: f.timestep_select(interval) for m, f in sims.items()}
    out = []
    try:
        for i in range(intervals + 1):
            fs = _state(sims[finest])
            out.append({m: l_comparison(_state(f), fs) for m, f in sims.items() if m != finest})
            if callback is not None:
                callback(i, out[-1])
            if i == intervals:
                break
            for m, f in sims.items():
                l, adt = sel[m]
                f.steps(adt, l)
    finally:
        for f in sims.values():
            f.close()
    return out, {m: s[0] for m, s in sel.items()}
