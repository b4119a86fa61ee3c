"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the reference's
level-set library compiled from its sources (oracle/_ref/libreflevel.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# INCRMT_ORACLE_SO points the tests at another build of the oracle (the
# planted-mistake builds of tests/test_oracle_pin_mutants.py)
ORACLE_SO = os.environ.get("INCRMT_ORACLE_SO") or os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libreflevel.so")
REF_INTERP_SO = os.path.join(ROOT, "oracle", "_ref", "librefinterp.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


_libs = {}


def load(path):
    if path not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run __graft_entry__.build())")
        _libs[path] = ctypes.CDLL(path)
    return _libs[path]


def oracle():
    return load(ORACLE_SO)


def reference():
    return load(REF_SO)


def have_reference():
    return os.path.exists(REF_SO)


def _band_call(lib, prefix, m, n, ax, ay, dx, dy, pd, pu, phi):
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    s = np.zeros(m * n, dtype=np.int32)
    order = np.zeros(m * n, dtype=np.int32)
    norder = ctypes.c_int(0)
    f = getattr(lib, prefix + "build_band")
    f.argtypes = [ctypes.c_int] * 2 + [ctypes.c_double] * 6 + [_dp, _ip, _ip, _ip]
    f(m, n, ax, ay, dx, dy, pd, pu, dptr(phi), iptr(s), iptr(order), ctypes.byref(norder))
    return phi, s, order[: norder.value].copy()


def build_band(which, m, n, ax, ay, dx, dy, pd, pu, phi):
    lib = oracle() if which == "oracle" else reference()
    prefix = "orc_" if which == "oracle" else "ref_"
    return _band_call(lib, prefix, m, n, ax, ay, dx, dy, pd, pu, phi)


def build_band_extrapolate(which, m, n, ax, ay, dx, dy, pd, pu, phi, fx, fy):
    lib = oracle() if which == "oracle" else reference()
    prefix = "orc_" if which == "oracle" else "ref_"
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    fx = np.ascontiguousarray(fx, dtype=np.float64).copy()
    fy = np.ascontiguousarray(fy, dtype=np.float64).copy()
    s = np.zeros(m * n, dtype=np.int32)
    f = getattr(lib, prefix + "build_band_extrapolate")
    f.argtypes = [ctypes.c_int] * 2 + [ctypes.c_double] * 6 + [_dp, _ip, _dp, _dp]
    rc = f(m, n, ax, ay, dx, dy, pd, pu, dptr(phi), iptr(s), dptr(fx), dptr(fy))
    return rc, phi, s, fx, fy


# ------------------------------------------------------------------ fluid
SPIN = 0
CIRCLE = 1
SQUARE = 2
FLAPPER = 3
LAMPREY = 4
FULL = 5


def simple_spin_spec(m, omega=1.0):
    """ftest simple-spin (ftest.cc:116-143): G=48, visc=1e-2, rhos=3,
    obj_simple_spin(0.8,0.5,0.125,0.2,pi,omega) in a no-slip [-1,1]^2 box."""
    import math
    return dict(m=m, n=m, ax=-1.0, bx=1.0, ay=-1.0, by=1.0, visc=1e-2, rhof=1.0,
                tw=2.5, pinw=2.5, w_scale=3.0, w_force=20.0, ev_mult=0.8, dt_pad=0.4,
                dt_ev_pad=0.8,
                objects=[(SPIN, [0.8, 0.5, 0.125, 0.2, math.pi, omega], 48.0, 3.0, 1.0)])


def conv_spec(m, ctype, coarsest):
    """conv_test's simulations (conv_test.cc:37-128) for test type ctype
    0..7 without the option bits: sim_velocity_pulses in the periodic
    [-1,1]^2 box, 0/1 fluid only, 3 obj_square(-0.1,0,0.6), 4..7
    obj_circle(-0.1,0,0.6) with the listed densities and viscosities, G=1,
    initialize(0.4, 0.4 coarsest/m, 0.8 coarsest/m); type 2 is obj_full
    with solid periodic boundaries (conv_test.cc:120)."""
    rhos = [1, 1, 1, 1, 3, 1, 1, 1][ctype]
    visc = [1e-3, 4e-3, 1e-20, 1e-3, 1e-3, 1e-3, 4e-3, 16e-3][ctype]
    kind = SQUARE if ctype == 3 else CIRCLE
    # conv_test's solids leave the pulses inside them (set_velocity false,
    # conv_test.cc:108)
    objs = [] if ctype in (0, 1) else [(kind, [-0.1, 0.0, 0.6, 0.0, 0.0, 0.0], 1.0, float(rhos), 1.0, 0)]
    if ctype == 2:
        objs = [(FULL, [0.0] * 9, 1.0, float(rhos), 1.0, 0)]
    return dict(m=m, n=m, ax=-1.0, bx=1.0, ay=-1.0, by=1.0, visc=visc, rhof=1.0, tw=2.5, pinw=2.5,
                w_scale=3.0, w_force=20.0, ev_mult=0.4, dt_pad=0.4 * coarsest / m,
                dt_ev_pad=0.8 * coarsest / m, x_prd=1, y_prd=1, vel=2, objects=objs,
                solid_prd_bc=int(ctype == 2))


def horiz_flow_spec(m, n, U=1.0, y_prd=True):
    """sim_horiz_flow(U, -1, 3, -1, 1) (sim_type.hh:116-145): inflow U from
    the left, zero pressure on the right edge, periodic in y (or between
    walls with y_prd False), with a soft circle off the mirror lines
    (obj_circle(-0.0937, 0.0311, 0.3, 0), G=2, rhos=1.5)."""
    return dict(m=m, n=n, ax=-1.0, bx=3.0, ay=-1.0, by=1.0, visc=1e-2, rhof=1.0, tw=2.5, pinw=2.5,
                w_scale=3.0, w_force=20.0, ev_mult=0.8, dt_pad=0.4, dt_ev_pad=0.8, x_prd=0,
                y_prd=int(y_prd), vel=0, zpr=1, U=float(U),
                objects=[(CIRCLE, [-0.0937, 0.0311, 0.3, 0.0, 0.0, 0.0], 2.0, 1.5, 1.0)])


class OracleFluid:
    def __init__(self, spec, mg_tol=1e-13):
        lib = oracle()
        self.lib = lib
        self.m, self.n = spec["m"], spec["n"]
        self.ml, self.nl = self.m + 4, self.n + 4
        objs = spec["objects"]
        kinds = np.array([o[0] for o in objs], dtype=np.int32)
        # objects: (kind, shape parameters, G, rhos, ev_trans_mult[, set_velocity])
        params = np.array([(list(o[1]) + [0.0] * 9)[:9] + [o[2], o[3], o[4], float(o[5] if len(o) > 5 else 1)]
                           for o in objs], dtype=np.float64).ravel()
        D, I = ctypes.c_double, ctypes.c_int
        lib.orc_fluid_create.restype = ctypes.c_void_p
        lib.orc_fluid_create.argtypes = [I, I] + [D] * 14 + [I, _ip, _dp, I, I, I, I, D, I]
        lib.orc_fluid_destroy.argtypes = [ctypes.c_void_p]
        lib.orc_fluid_step.argtypes = [ctypes.c_void_p, D]
        lib.orc_fluid_dt_reg.argtypes = [ctypes.c_void_p]
        lib.orc_fluid_dt_reg.restype = D
        lib.orc_fluid_time.argtypes = [ctypes.c_void_p]
        lib.orc_fluid_time.restype = D
        lib.orc_fluid_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, I, _dp]
        lib.orc_fluid_set.argtypes = [ctypes.c_void_p, ctypes.c_char_p, I, _dp]
        lib.orc_fluid_stats.argtypes = [ctypes.c_void_p, _dp]
        self.h = lib.orc_fluid_create(
            spec["m"], spec["n"], spec["ax"], spec["bx"], spec["ay"], spec["by"], spec["visc"],
            spec["rhof"], spec["tw"], spec["pinw"], spec["w_scale"], spec["w_force"],
            spec["ev_mult"], spec["dt_pad"], spec["dt_ev_pad"], mg_tol, len(objs),
            iptr(kinds), dptr(params), int(spec.get("x_prd", 0)), int(spec.get("y_prd", 0)),
            int(spec.get("vel", 0)), int(spec.get("zpr", 0)), float(spec.get("U", 0.0)),
            int(spec.get("solid_prd_bc", 0)))
        self.nobj = len(objs)
        self.x_prd, self.y_prd = int(spec.get("x_prd", 0)), int(spec.get("y_prd", 0))
        dx, dy = (spec["bx"] - spec["ax"]) / self.m, (spec["by"] - spec["ay"]) / self.n
        self.eps = spec["tw"] * (dx if dx < dy else dy)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_fluid_destroy(self.h)
            self.h = None

    @property
    def dt_reg(self):
        return self.lib.orc_fluid_dt_reg(self.h)

    @property
    def time(self):
        return self.lib.orc_fluid_time(self.h)

    def step(self, dt):
        rc = self.lib.orc_fluid_step(self.h, dt)
        if rc:
            raise RuntimeError(f"oracle step failed with code {rc}")

    def get(self, name, k=0):
        out = np.zeros(self.ml * self.nl)
        if self.lib.orc_fluid_get(self.h, name.encode(), k, dptr(out)):
            raise KeyError(name)
        return out.reshape(self.nl, self.ml)

    def set(self, name, arr, k=0):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if self.lib.orc_fluid_set(self.h, name.encode(), k, dptr(a)):
            raise KeyError(name)

    def phase(self, name, dt=0.0, zfem=None):
        """One phase of the step on the current state (oracle fluid_phase):
        returns the FEM source for "fem_source"; "copy_pressure" takes the
        FEM solution zfem (m_fem x n_fem)."""
        mf = self.m if getattr(self, "x_prd", 0) else self.m + 1
        nf = self.n if getattr(self, "y_prd", 0) else self.n + 1
        out = np.zeros(mf * nf)
        zin = np.zeros(mf * nf) if zfem is None else np.ascontiguousarray(zfem, dtype=np.float64).ravel()
        f = self.lib.orc_fluid_phase
        f.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double, _dp, _dp]
        if f(self.h, name.encode(), dt, dptr(zin), dptr(out)):
            raise KeyError(name)
        return out.reshape(nf, mf)

    def set_time(self, t):
        self.lib.orc_fluid_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double]
        self.lib.orc_fluid_set_time(self.h, t)

    def stats(self):
        out = np.zeros(4)
        self.lib.orc_fluid_stats(self.h, dptr(out))
        return dict(mac_cycles=int(out[0]), mac_res=out[1], fem_cycles=int(out[2]), fem_res=out[3])

    # tracers (fluid_2d.cc:1376-1484, object.cc:313-417)
    def init_tracers(self, ntrace):
        self.lib.orc_fluid_init_tracers.argtypes = [ctypes.c_void_p, ctypes.c_int]
        self.lib.orc_fluid_init_tracers(self.h, ntrace)

    def frame_output(self):
        self.lib.orc_fluid_frame_output.argtypes = [ctypes.c_void_p]
        self.lib.orc_fluid_frame_output(self.h)

    def tracers(self):
        self.lib.orc_fluid_ntrace.argtypes = [ctypes.c_void_p]
        nt = self.lib.orc_fluid_ntrace(self.h)
        out = np.zeros(4 * nt)
        self.lib.orc_fluid_get_tracers.argtypes = [ctypes.c_void_p, _dp]
        self.lib.orc_fluid_get_tracers(self.h, dptr(out))
        return out.reshape(nt, 4)

    def set_tracers(self, tr):
        a = np.ascontiguousarray(tr, dtype=np.float64).ravel()
        self.lib.orc_fluid_set_tracers.argtypes = [ctypes.c_void_p, _dp]
        self.lib.orc_fluid_set_tracers(self.h, dptr(a))

    def track_state(self, k=0):
        out = np.zeros(10)
        self.lib.orc_fluid_get_track.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        if self.lib.orc_fluid_get_track(self.h, k, dptr(out)):
            raise KeyError(k)
        return out

    def set_track_state(self, st, k=0):
        a = np.ascontiguousarray(st, dtype=np.float64)
        self.lib.orc_fluid_set_track.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        if self.lib.orc_fluid_set_track(self.h, k, dptr(a)):
            raise KeyError(k)

    def track(self, k=0):
        """track.dat records of object k: (time, x, y, unwound angle)."""
        f = self.lib.orc_fluid_track_records
        f.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int]
        n = f(self.h, k, dptr(np.zeros(4)), 0)
        out = np.zeros(4 * max(n, 1))
        f(self.h, k, dptr(out), n)
        return out[: 4 * n].reshape(n, 4)


    def timestep_select(self, interval):
        return timestep_select(interval, self.dt_reg)

    def steps(self, dt, k):
        for _ in range(k):
            self.step(dt)

    def close(self):
        pass


def timestep_select(interval, dt_reg):
    """fluid_2d::timestep_select (fluid_2d.hh:173-177)."""
    l = int(interval / dt_reg) + 1
    return l, interval / l


def interp2(which, m, n, ax, bx, ay, by, fx, fy, x, y, solid=True):
    """bicubic_interp (bi_interp.cc) of (fx, fy) at (x, y): value, d/dx, d/dy.
    which: "oracle" or "reference" (the reference's bi_interp.cc compiled
    from its sources)."""
    fx = np.ascontiguousarray(fx, dtype=np.float64).ravel()
    fy = np.ascontiguousarray(fy, dtype=np.float64).ravel()
    out = np.zeros(6)
    D = ctypes.c_double
    if which == "oracle":
        f = oracle().orc_interp2
        f.argtypes = [ctypes.c_int] * 2 + [D] * 4 + [_dp, _dp, D, D, _dp]
        f(m, n, ax, bx, ay, by, dptr(fx), dptr(fy), x, y, dptr(out))
        if not solid:
            out[2:] = 0
    else:
        f = load(REF_INTERP_SO).ref_interp2
        f.argtypes = [ctypes.c_int] * 2 + [D] * 4 + [_dp, _dp, D, D, ctypes.c_int, _dp]
        f(m, n, ax, bx, ay, by, dptr(fx), dptr(fy), x, y, int(solid), dptr(out))
    return out


def glibc_rand_sequence(n):
    out = np.zeros(n, dtype=np.int32)
    f = oracle().orc_glibc_rand
    f.argtypes = [ctypes.c_int, _ip]
    f(n, iptr(out))
    return out
