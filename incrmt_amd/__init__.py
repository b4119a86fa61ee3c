"""MI355X IncRMT hot path: Python front end of libincrmt_hip.so.

The classes mirror the reference's interfaces for the per-timestep path:

* ``Fluid2D``   -- fluid_2d (fluid_2d.hh:31-272): add_object, initialize,
  step_forward, timestep_select, solve; state access on the reference's
  ghosted (m+4)x(n+4) grid.
* ``MgSolver``  -- the mgs_mac_* / mgs_fem_* projection solvers
  (mgs_mac.hh:41-102, mgs_fem.hh:61-136) with their standalone constructor.
* ``build_band`` / ``extrapolate`` -- levelset::build_band and
  levelset::extrapolate_fields (levelset.cc:199-335, traverse.hh:235-260).

Everything runs in the HIP library; there is no CPU fallback. Importing this
module does not touch the GPU, but any call fails with ``IncrmtError`` when
the library has not been built.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libincrmt_hip.so")

OK = 0
ERR_ARG, ERR_NOT_CONVERGED, ERR_NEWTON, ERR_BAND, ERR_SINGULAR_FIT, ERR_HIP, ERR_UNSUPPORTED = range(1, 8)
MG_MAC, MG_FEM = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
GUESS_PREVIOUS, GUESS_EXTRAPOLATE, GUESS_QUADRATIC, GUESS_LEARNED = 0, 1, 2, 3
LAUNCH_DIRECT, LAUNCH_GRAPH = 0, 1
SIM_REST, SIM_FOUR_ROLLS, SIM_VELOCITY_PULSES = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_D, _I, _V = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)


class IncrmtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"incrmt error {code}: {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [("mac_cycles", _I), ("fem_cycles", _I), ("mac_rel_res", _D), ("fem_rel_res", _D),
                ("total_mac_cycles", ctypes.c_long), ("total_fem_cycles", ctypes.c_long),
                ("steps", ctypes.c_long), ("reinit_passes", _I), ("extend_passes", _I),
                ("smoother_ms", _D), ("smoother_launches", ctypes.c_long), ("smoother_bytes", _D),
                ("proj_ms", _D)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) of every entry point in include/incrmt.h
_PROTOS = [
    ("incrmt_last_error", ctypes.c_char_p, []),
    ("incrmt_device_info", _I, [ctypes.c_char_p, _I, _dp]),
    ("incrmt_copy_probe", _I, [ctypes.c_longlong, _I, _dp]),
    ("incrmt_fluid_create", _I, [_I, _I, _D, _D, _D, _D, _I, _I, _D, _D, _D, _D, _D, _D, ctypes.POINTER(_V)]),
    ("incrmt_fluid_add_spin", _I, [_V, _D, _D, _D, _D, _D, _D, _D, _D, _D]),
    ("incrmt_fluid_add_circle", _I, [_V, _D, _D, _D, _D, _D, _D, _D]),
    ("incrmt_fluid_add_square", _I, [_V, _D, _D, _D, _D, _D, _D, _D]),
    ("incrmt_fluid_add_flapper", _I, [_V] + [_D] * 12),
    ("incrmt_fluid_set_solid_velocity", _I, [_V, _I, _I]),
    ("incrmt_fluid_add_full", _I, [_V, _D, _D, _D]),
    ("incrmt_fluid_set_solid_periodic", _I, [_V, _I]),
    ("incrmt_fluid_add_lamprey", _I, [_V] + [_D] * 12),
    ("incrmt_fluid_initialize", _I, [_V, _D, _D, _D]),
    ("incrmt_fluid_set_solver", _I, [_V, _D, _I, _I]),
    ("incrmt_fluid_step", _I, [_V, _D]),
    ("incrmt_fluid_steps", _I, [_V, _D, _I]),
    ("incrmt_fluid_dt_reg", _D, [_V]),
    ("incrmt_fluid_time", _D, [_V]),
    ("incrmt_fluid_timestep_select", _I, [_V, _D, _dp]),
    ("incrmt_fluid_download", _I, [_V, ctypes.c_char_p, _I, _dp]),
    ("incrmt_fluid_upload", _I, [_V, ctypes.c_char_p, _I, _dp]),
    ("incrmt_fluid_set_time", _I, [_V, _D]),
    ("incrmt_fluid_sync", _I, [_V]),
    ("incrmt_fluid_get_stats", _I, [_V, ctypes.POINTER(Stats)]),
    ("incrmt_fluid_time_smoothers", _I, [_V, _I]),
    ("incrmt_fluid_set_solid_windows", _I, [_V, _I]),
    ("incrmt_fluid_destroy", None, [_V]),
    ("incrmt_fluid_set_ntrace", _I, [_V, _I]),
    ("incrmt_fluid_describe", _I, [_V, ctypes.c_char_p, _I]),
    ("incrmt_fluid_set_guess", _I, [_V, _I]),
    ("incrmt_fluid_set_launch", _I, [_V, _I]),
    ("incrmt_fluid_set_overlap", _I, [_V, _I]),
    ("incrmt_fluid_set_sim_velocity", _I, [_V, _I]),
    ("incrmt_fluid_set_horizontal_flow", _I, [_V, _D]),
    ("incrmt_fluid_time_fine_smoother", _I, [_V, _I, _I, _I, _dp, _ip]),
    ("incrmt_fluid_write_frame", _I, [_V, ctypes.c_char_p, _I, ctypes.c_uint]),
    ("incrmt_fluid_solve", _I, [_V, ctypes.c_char_p, _D, _I, ctypes.c_uint]),
    ("incrmt_fluid_get_tracers", _I, [_V, _dp]),
    ("incrmt_fluid_set_tracers", _I, [_V, _dp]),
    ("incrmt_fluid_get_track_state", _I, [_V, _I, _dp]),
    ("incrmt_fluid_set_track_state", _I, [_V, _I, _dp]),
    ("incrmt_fluid_frame_output", _I, [_V]),
    ("incrmt_fluid_track", _I, [_V, _I, _dp, _I, _ip]),
    ("incrmt_mg_create", _I, [_I, _I, _I, _I, _I, _D, _D, ctypes.POINTER(_V)]),
    ("incrmt_mg_create_bc", _I, [_I, _I, _I, _I, _I, _I, _D, _D, ctypes.POINTER(_V)]),
    ("incrmt_mg_set_irho", _I, [_V, _dp]),
    ("incrmt_mg_solve", _I, [_V, _dp, _dp, _D, _I, _I, _ip, _dp]),
    ("incrmt_mg_time_solve", _I, [_V, _D, _I, _I, _I, _dp, _ip, _dp]),
    ("incrmt_mg_apply", _I, [_V, _dp, _dp]),
    ("incrmt_mg_smooth", _I, [_V, _dp, _dp, _I, _I]),
    ("incrmt_mg_smooth_prec", _I, [_V, _dp, _dp, _I, _I, _I]),
    ("incrmt_mg_set_launch", _I, [_V, _I]),
    ("incrmt_mg_set_krylov", _I, [_V, _I]),
    ("incrmt_fluid_set_krylov", _I, [_V, _I]),
    ("incrmt_mg_time_smoother", _I, [_V, _I, _I, _dp, _ip]),
    ("incrmt_mg_time_shape", _I, [_V, _I, _I, _I, _dp, _ip]),
    ("incrmt_mg_destroy", None, [_V]),
    ("incrmt_levelset_build_band", _I, [_I, _I, _D, _D, _D, _D, _D, _D, _dp, _ip, _ip]),
    ("incrmt_levelset_build_band_noalloc", _I, [_I, _I, _D, _D, _D, _D, _D, _D, _V, _V, _ip]),
    ("incrmt_fluid_set_guess_training", _I, [_V, _I, ctypes.c_float, _D]),
    ("incrmt_fluid_time_projections", _I, [_V, _I]),
    ("incrmt_fluid_guess_nparams", _I, [_V, _I, ctypes.POINTER(ctypes.c_longlong)]),
    ("incrmt_fluid_get_guess_params", _I, [_V, _I, _fp]),
    ("incrmt_fluid_set_guess_params", _I, [_V, _I, _fp]),
    ("incrmt_fluid_guess_stats", _I, [_V, _I, _dp]),
    ("incrmt_fluid_set_guess_cost", _I, [_V, _I, _D]),
    ("incrmt_fluid_guess_cost_stats", _I, [_V, _I, _dp]),
    ("incrmt_guess_policy_run", _I, [_dp, _I, _D, _ip]),
    ("incrmt_unet_create", _I, [_I, _I, ctypes.c_uint, ctypes.POINTER(_V)]),
    ("incrmt_unet_nparams", _I, [_V, ctypes.POINTER(ctypes.c_longlong)]),
    ("incrmt_unet_get_params", _I, [_V, _fp]),
    ("incrmt_unet_set_params", _I, [_V, _fp]),
    ("incrmt_unet_forward", _I, [_V, _fp, _fp]),
    ("incrmt_unet_gradient", _I, [_V, _fp, _fp, _dp, _fp]),
    ("incrmt_unet_train_step", _I, [_V, _fp, _fp, ctypes.c_float, _dp]),
    ("incrmt_unet_destroy", None, [_V]),
    ("incrmt_levelset_extrapolate", _I, [_I, _I, _D, _D, _D, _D, _D, _D, _dp, _ip, _dp, _dp, _ip]),
]
SYMBOLS = [p[0] for p in _PROTOS]

_lib = None


def lib():
    """The loaded HIP library; raises IncrmtError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise IncrmtError(ERR_HIP, f"{LIB_PATH} is missing: build it with "
                              "`python __graft_entry__.py` or incrmt_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _PROTOS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        raise IncrmtError(rc, lib().incrmt_last_error().decode())


def _d(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def copy_probe(nbytes, reps=10):
    """Milliseconds per device-to-device copy of nbytes (HIP events)."""
    ms = _D(0)
    _check(lib().incrmt_copy_probe(nbytes, reps, ctypes.byref(ms)))
    return ms.value


def device_info():
    name = ctypes.create_string_buffer(256)
    hbm = _D(0)
    _check(lib().incrmt_device_info(name, 256, ctypes.byref(hbm)))
    return name.value.decode(), hbm.value


class Fluid2D:
    """fluid_2d (fluid_2d.hh:31-272) on one GPU: a box with no-slip walls or
    periodic directions (x_prd, y_prd; even cell counts whose halvings end at
    2 or 3)."""

    def __init__(self, m, n, visc, rhof, tw=2.5, pinw=2.5, ax=-1.0, bx=1.0, ay=-1.0, by=1.0,
                 x_prd=False, y_prd=False, w_scale=3.0, w_force=20.0):
        L = lib()
        h = _V()
        _check(L.incrmt_fluid_create(m, n, ax, bx, ay, by, int(x_prd), int(y_prd), w_scale, w_force,
                                     visc, rhof, tw, pinw, ctypes.byref(h)))
        self._h = h
        self.m, self.n = m, n
        self.ml, self.nl = m + 4, n + 4
        self.nobj = 0
        dx, dy = (bx - ax) / m, (by - ay) / n
        self.eps = tw * (dx if dx < dy else dy)  # the transition width (fluid_2d.cc:35)
        self.ntrace = 0
        self.x_prd, self.y_prd = bool(x_prd), bool(y_prd)

    def close(self):
        if getattr(self, "_h", None):
            lib().incrmt_fluid_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # obj_simple_spin / obj_circle / obj_square with mat_const (object.hh:128-150, 422-467)
    def add_spin(self, r_scale, a1, a2, piv_r, th_max, omega, G, rhos, ev_trans_mult=1.0):
        _check(lib().incrmt_fluid_add_spin(self._h, r_scale, a1, a2, piv_r, th_max, omega, G, rhos, ev_trans_mult))
        self.nobj += 1

    def add_circle(self, cx, cy, cr, grav, G, rhos, ev_trans_mult=1.0):
        _check(lib().incrmt_fluid_add_circle(self._h, cx, cy, cr, grav, G, rhos, ev_trans_mult))
        self.nobj += 1

    def add_square(self, cx, cy, cr, grav, G, rhos, ev_trans_mult=1.0):
        _check(lib().incrmt_fluid_add_square(self._h, cx, cy, cr, grav, G, rhos, ev_trans_mult))
        self.nobj += 1

    def add_full(self, G, rhos, ev_trans_mult=1.0):
        """obj_full (object.hh:83-86): a solid filling the box."""
        _check(lib().incrmt_fluid_add_full(self._h, G, rhos, ev_trans_mult))
        self.nobj += 1

    def set_solid_periodic(self, on=True):
        """fluid_2d::solid_prd_bc: periodic reference maps shifted by the box
        lengths, no wall forces on solids; doubly periodic boxes only."""
        _check(lib().incrmt_fluid_set_solid_periodic(self._h, int(bool(on))))

    def set_solid_velocity(self, k, set_velocity):
        """mat_const::set_velocity of solid k (obj_field.hh:25-28): False
        leaves the sim_type's initial velocity inside it (conv_test)."""
        _check(lib().incrmt_fluid_set_solid_velocity(self._h, int(k), int(bool(set_velocity))))

    def add_flapper(self, cx, cy, cr, cl, ar, al, theta, m_stretch, T, G, rhos, ev_trans_mult=1.0):
        """obj_flapper(cx,cy,cr,cl,ar,al,theta,m_stretch,T) (object.hh:245-289)."""
        _check(lib().incrmt_fluid_add_flapper(self._h, cx, cy, cr, cl, ar, al, theta, m_stretch, T, G, rhos,
                                              ev_trans_mult))
        self.nobj += 1

    def add_lamprey(self, cx, cy, theta, le, dis, w1, w2, T, mlam, G, rhos, ev_trans_mult=1.0):
        """obj_lamprey(cx,cy,theta,le,dis,w1,w2,T,mlam) (object.hh:542-595)."""
        _check(lib().incrmt_fluid_add_lamprey(self._h, cx, cy, theta, le, dis, w1, w2, T, mlam, G, rhos,
                                              ev_trans_mult))
        self.nobj += 1

    def initialize(self, ev_mult=0.8, dt_pad=0.4, dt_ev_pad=0.8):
        _check(lib().incrmt_fluid_initialize(self._h, ev_mult, dt_pad, dt_ev_pad))

    def set_solver(self, rel_tol=1e-8, max_cycles=100, precision=PREC_FP64):
        _check(lib().incrmt_fluid_set_solver(self._h, rel_tol, max_cycles, precision))

    def set_guess(self, mode):
        """The projections' starting iterate: GUESS_PREVIOUS (the
        reference's warm start, the default), GUESS_EXTRAPOLATE (the linear
        extrapolation in time of the last two solutions,
        mgs_base::initial_guess with mg_lin_t_interp, mgs_common.cc:33-66)
        or GUESS_QUADRATIC (the quadratic extrapolation through the last
        three). Extrapolated starts are made mean-free for the singular
        (pure Neumann) systems; changing the mode clears the history."""
        _check(lib().incrmt_fluid_set_guess(self._h, mode))

    def set_guess_training(self, train=True, lr=1e-3, fallback_tol=1e-4):
        """Online training of the two learned-guess networks: one Adam step
        per converged projection (the loss is the relative energy norm of
        the error left after the learned start's step along the
        prediction);
        fallback_tol is the relative residual after one V-cycle above which
        a learned start is dropped for the quadratic start and FP64 cycles."""
        _check(lib().incrmt_fluid_set_guess_training(self._h, int(train), lr, fallback_tol))

    def set_guess_cost(self, mode=1, cycles=0.0):
        """Use the learned start only where it pays: mode 1 measures its
        cost in V-cycles, mode 2 takes `cycles`, mode 0 always uses it."""
        _check(lib().incrmt_fluid_set_guess_cost(self._h, mode, cycles))

    def guess_cost_stats(self, kind):
        o = (ctypes.c_double * 6)()
        _check(lib().incrmt_fluid_guess_cost_stats(self._h, kind, o))
        return dict(cost=o[0], mean_learned=o[1], mean_control=o[2], off_for_cost=bool(o[3]),
                    learned_counted=int(o[4]), control_counted=int(o[5]))

    def time_projections(self, enable=True):
        """HIP events around each step's two projections (stats proj_ms)."""
        _check(lib().incrmt_fluid_time_projections(self._h, int(enable)))

    def guess_params(self, kind):
        n = ctypes.c_longlong(0)
        _check(lib().incrmt_fluid_guess_nparams(self._h, kind, ctypes.byref(n)))
        p = np.zeros(n.value, dtype=np.float32)
        _check(lib().incrmt_fluid_get_guess_params(self._h, kind, _f(p)))
        return p

    def set_guess_params(self, kind, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().incrmt_fluid_set_guess_params(self._h, kind, _f(p)))

    def guess_stats(self, kind):
        o = np.zeros(8)
        _check(lib().incrmt_fluid_guess_stats(self._h, kind, _d(o)))
        return dict(learned_solves=int(o[0]), fallbacks=int(o[1]), train_steps=int(o[2]), loss=o[3],
                    off=bool(o[4]), misses=int(o[5]), rel0=o[6], rel1=o[7])

    def set_sim_velocity(self, kind):
        """SIM_REST, SIM_FOUR_ROLLS or SIM_VELOCITY_PULSES: the sim_type's
        initial fluid velocity (sim_type.cc:30-47); before initialize."""
        _check(lib().incrmt_fluid_set_sim_velocity(self._h, kind))

    def set_horizontal_flow(self, U):
        """sim_horiz_flow (sim_type.hh:116-145): initial velocity (U, 0),
        inflow (U, 0) through the left ghost columns and zero pressure on the
        right edge (fluid_2d::zpr); x must not be periodic. Before
        initialize."""
        _check(lib().incrmt_fluid_set_horizontal_flow(self._h, float(U)))

    def set_launch(self, mode):
        """LAUNCH_GRAPH (V-cycles captured once as HIP graphs and replayed)
        or LAUNCH_DIRECT (kernel by kernel); bit-identical. Without a call
        the solvers pick by size: graphs below 512^2 unknowns, direct
        launches above. A call before initialize is kept and applied there."""
        _check(lib().incrmt_fluid_set_launch(self._h, mode))

    def set_krylov(self, on):
        """Both projections' mixed-precision solves as FP64 conjugate
        gradients preconditioned by one FP32 V-cycle per iteration."""
        _check(lib().incrmt_fluid_set_krylov(self._h, int(bool(on))))

    def set_overlap(self, on):
        """Run the end-of-step band rebuild (a lone solid's, or the shared
        one of several solids) on a second stream beside the FEM projection
        (True, the default) or after it (False); bit-identical."""
        _check(lib().incrmt_fluid_set_overlap(self._h, int(bool(on))))

    def step_forward(self, dt):
        _check(lib().incrmt_fluid_step(self._h, dt))

    def steps(self, dt, k):
        _check(lib().incrmt_fluid_steps(self._h, dt, k))

    def sync(self):
        _check(lib().incrmt_fluid_sync(self._h))

    @property
    def dt_reg(self):
        return lib().incrmt_fluid_dt_reg(self._h)

    @property
    def time(self):
        return lib().incrmt_fluid_time(self._h)

    @time.setter
    def time(self, t):
        _check(lib().incrmt_fluid_set_time(self._h, t))

    def timestep_select(self, interval):
        adt = _D(0)
        l = lib().incrmt_fluid_timestep_select(self._h, interval, ctypes.byref(adt))
        if l < 0:
            raise IncrmtError(ERR_ARG, "timestep_select before initialize")
        return l, adt.value

    def solve(self, duration, frames, callback=None):
        """fluid_2d::solve (fluid_2d.cc:260-294) without file output."""
        l, adt = self.timestep_select(duration / frames)
        for k in range(1, frames + 1):
            self.steps(adt, l)
            if callback is not None:
                callback(k, self)

    def get(self, name, k=0):
        out = np.zeros(self.ml * self.nl)
        _check(lib().incrmt_fluid_download(self._h, name.encode(), k, _d(out)))
        return out.reshape(self.nl, self.ml)

    def set(self, name, arr, k=0):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if a.size != self.ml * self.nl:
            raise IncrmtError(ERR_ARG, "field has the wrong size")
        _check(lib().incrmt_fluid_upload(self._h, name.encode(), k, _d(a)))

    def stats(self):
        s = Stats()
        _check(lib().incrmt_fluid_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def time_smoothers(self, enable=True):
        _check(lib().incrmt_fluid_time_smoothers(self._h, int(enable)))

    def set_solid_windows(self, on=True):
        """Per-solid sweeps over each solid's band box (default) or the
        whole grid; the fields are the same bit for bit. Set before
        initialize to have no windowed step at all."""
        _check(lib().incrmt_fluid_set_solid_windows(self._h, int(on)))

    def write_frame(self, directory, sn, fflags=8 | 16 | 64 | 128 | 256 | 512):
        """fluid_2d::write_files(sn): the fflags-selected fields as float32
        Gnuplot matrices in `directory` (ftest's default flags)."""
        _check(lib().incrmt_fluid_write_frame(self._h, str(directory).encode(), sn, fflags))

    def solve(self, duration, frames, directory=None, fflags=8 | 16 | 64 | 128 | 256 | 512):
        """fluid_2d::solve(duration, frames): `frames` frames of
        timestep_select's steps, each written to `directory` (when given)
        and reported by one line on the process's stdout (machine-readable
        with fflags & 8192)."""
        d = None if directory is None else str(directory).encode()
        _check(lib().incrmt_fluid_solve(self._h, d, duration, frames, fflags))

    def time_fine_smoother(self, kind, sweeps, precision=PREC_FP64):
        """Back-to-back fine-level sweeps of one projection system between two
        HIP events: (ms per launch, launches)."""
        ms, n = _D(0), _I(0)
        _check(lib().incrmt_fluid_time_fine_smoother(self._h, kind, sweeps, precision, ctypes.byref(ms),
                                                     ctypes.byref(n)))
        return ms.value, n.value

    def describe(self):
        """The autotuned fused-smoother shapes per multigrid level."""
        buf = ctypes.create_string_buffer(2048)
        _check(lib().incrmt_fluid_describe(self._h, buf, 2048))
        return buf.value.decode()

    # tracers (fluid_2d.cc:1376-1484, object.cc:202-219, 313-417)
    def set_ntrace(self, ntrace):
        """fluid_2d's ntrace: passive tracers placed by initialize."""
        _check(lib().incrmt_fluid_set_ntrace(self._h, ntrace))
        self.ntrace = ntrace

    def tracers(self):
        out = np.zeros(4 * self.ntrace)
        _check(lib().incrmt_fluid_get_tracers(self._h, _d(out)))
        return out.reshape(self.ntrace, 4)

    def set_tracers(self, tr):
        a = np.ascontiguousarray(tr, dtype=np.float64).ravel()
        if a.size != 4 * self.ntrace:
            raise IncrmtError(ERR_ARG, "tracer array has the wrong size")
        _check(lib().incrmt_fluid_set_tracers(self._h, _d(a)))

    def track_state(self, k=0):
        out = np.zeros(10)
        _check(lib().incrmt_fluid_get_track_state(self._h, k, _d(out)))
        return out

    def set_track_state(self, st, k=0):
        a = np.ascontiguousarray(st, dtype=np.float64)
        _check(lib().incrmt_fluid_set_track_state(self._h, k, _d(a)))

    def frame_output(self):
        """The object part of fluid_2d::write_files: one track.dat record."""
        _check(lib().incrmt_fluid_frame_output(self._h))

    def track(self, k=0):
        """track.dat records of solid k: (time, x, y, unwound angle)."""
        n = ctypes.c_int(0)
        _check(lib().incrmt_fluid_track(self._h, k, None, 0, ctypes.byref(n)))
        cap = min(n.value, 16384)
        out = np.zeros(4 * max(cap, 1))
        _check(lib().incrmt_fluid_track(self._h, k, _d(out), cap, ctypes.byref(n)))
        return out[: 4 * cap].reshape(cap, 4)


def simple_spin(m, omega=1.0, rel_tol=1e-8, precision=PREC_FP64, guess=GUESS_PREVIOUS):
    """`ftest simple-spin m` (ftest.cc:116-143): G=48, visc=1e-2, rhos=3, the
    three-pronged rotor obj_simple_spin(0.8,0.5,0.125,0.2,pi,omega) in a
    no-slip [-1,1]^2 box, initialized with ev_mult 0.8, dt_pad 0.4,
    dt_ev_pad 0.8 (ftest.cc:100-103)."""
    f = Fluid2D(m, m, visc=1e-2, rhof=1.0, tw=2.5, pinw=2.5)
    f.add_spin(0.8, 0.5, 0.125, 0.2, math.pi, omega, G=48.0, rhos=3.0, ev_trans_mult=1.0)
    f.set_solver(rel_tol=rel_tol, precision=precision)
    f.set_guess(guess)
    f.initialize(0.8, 0.4, 0.8)
    return f


class MgSolver:
    """mgs_mac_* / mgs_fem_* with the standalone constructor
    (mgs_mac.cc:20-25, mgs_fem.cc:39-43): m x n cells (MAC) or nodes (FEM).
    zpr: zero pressure on the right edge (mgs_base's zpr,
    mgs_common.cc:15-17); a FEM system then keeps its fixed node column
    m-1 in every array, read and written as zero."""

    def __init__(self, kind, m, n, dx, dy, x_prd=False, y_prd=False, zpr=False):
        h = _V()
        _check(lib().incrmt_mg_create_bc(kind, m, n, int(x_prd), int(y_prd), int(zpr), dx, dy, ctypes.byref(h)))
        self._h = h
        self.kind, self.m, self.n = kind, m, n
        self.z = np.zeros(m * n)

    def close(self):
        if getattr(self, "_h", None):
            lib().incrmt_mg_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_irho(self, irho):
        if irho is None:
            _check(lib().incrmt_mg_set_irho(self._h, None))
        else:
            a = np.ascontiguousarray(irho, dtype=np.float64)
            self._irho = a
            _check(lib().incrmt_mg_set_irho(self._h, _d(a)))

    def solve(self, b, rel_tol=1e-10, max_cycles=200, precision=PREC_FP64, z=None):
        """Solve A z = b; z warm-starts from the previous solution
        (mgs_common.hh:48-50). Returns (z, cycles, relative residual)."""
        bb = np.ascontiguousarray(b, dtype=np.float64).ravel()
        if z is not None:
            self.z = np.ascontiguousarray(z, dtype=np.float64).ravel().copy()
        c, r = _I(0), _D(0)
        _check(lib().incrmt_mg_solve(self._h, _d(bb), _d(self.z), rel_tol, max_cycles, precision,
                                     ctypes.byref(c), ctypes.byref(r)))
        return self.z.copy(), c.value, r.value

    def time_solve(self, rel_tol=1e-10, max_cycles=200, precision=PREC_FP64, reps=3):
        """Device time of solves of the last solve's right-hand side (kept
        on the device) from a zero start: (ms per solve, cycles, relative
        residual)."""
        ms, c, r = _D(0), _I(0), _D(0)
        _check(lib().incrmt_mg_time_solve(self._h, rel_tol, max_cycles, precision, reps, ctypes.byref(ms),
                                          ctypes.byref(c), ctypes.byref(r)))
        return ms.value, c.value, r.value

    def set_launch(self, mode):
        """LAUNCH_GRAPH or LAUNCH_DIRECT V-cycles (default by size: graphs
        below 512^2 unknowns); bit-identical."""
        _check(lib().incrmt_mg_set_launch(self._h, mode))

    def set_krylov(self, on):
        """Mixed-precision solves as FP64 conjugate gradients preconditioned
        by one FP32 V-cycle per iteration (walled systems without a
        zero-pressure edge), instead of V-cycle refinement."""
        _check(lib().incrmt_mg_set_krylov(self._h, int(bool(on))))

    def apply(self, x):
        xx = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.zeros_like(xx)
        _check(lib().incrmt_mg_apply(self._h, _d(xx), _d(y)))
        return y

    def smooth(self, b, z, sweeps, shape, precision=PREC_FP64):
        """`sweeps` Gauss-Seidel sweeps from z with fused launch shape
        `shape`, or one launch per colour (shape -1); PREC_MIXED sweeps in
        single precision. Returns the new z."""
        bb = np.ascontiguousarray(b, dtype=np.float64).ravel()
        zz = np.ascontiguousarray(z, dtype=np.float64).ravel().copy()
        _check(lib().incrmt_mg_smooth_prec(self._h, _d(bb), _d(zz), sweeps, shape, precision))
        return zz

    def time_smoother(self, sweeps, precision=PREC_FP64, shape=-1):
        """Mean ms per fine-level smoother launch over `sweeps` sweeps and
        the launch count, on the autotuned shape or on `shape`."""
        ms, n = _D(0), _I(0)
        _check(lib().incrmt_mg_time_shape(self._h, sweeps, precision, shape, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


def build_band(m, n, ax, ay, dx, dy, phi_down, phi_up, phi):
    """levelset::build_band on an m x n grid. Returns (phi, status, passes)."""
    p = np.ascontiguousarray(phi, dtype=np.float64).ravel().copy()
    s = np.zeros(m * n, dtype=np.int32)
    passes = _I(0)
    _check(lib().incrmt_levelset_build_band(m, n, ax, ay, dx, dy, phi_down, phi_up, _d(p), _i(s),
                                            ctypes.byref(passes)))
    return p, s, passes.value


def extrapolate(m, n, ax, ay, dx, dy, phi_down, phi_up, phi, fx, fy):
    """build_band then levelset::extrapolate_fields of (fx, fy).
    Returns (phi, status, fx, fy, extension passes)."""
    p = np.ascontiguousarray(phi, dtype=np.float64).ravel().copy()
    a = np.ascontiguousarray(fx, dtype=np.float64).ravel().copy()
    b = np.ascontiguousarray(fy, dtype=np.float64).ravel().copy()
    s = np.zeros(m * n, dtype=np.int32)
    passes = _I(0)
    _check(lib().incrmt_levelset_extrapolate(m, n, ax, ay, dx, dy, phi_down, phi_up, _d(p), _i(s), _d(a), _d(b),
                                             ctypes.byref(passes)))
    return p, s, a, b, passes.value


def multibody(m=2048, count=64, seed=20260102, rel_tol=1e-8, precision=PREC_MIXED, max_cycles=100):
    """The sediment-style multibody configuration: `count` obj_circle solids
    (object.hh:128-138) placed by workloads.sediment_circles in a no-slip
    [-1,1]^2 box at m x m cells; solid/fluid density ratio 2, shear modulus
    1, viscosity 0.01, gravity 1 (obj_circle's accel, object.cc:31-34);
    transition and pinning widths 2.5 cells, ev_mult 0.8, dt_pad 0.4,
    dt_ev_pad 0.8 (the sediment.cc settings)."""
    from .workloads import sediment_circles
    f = Fluid2D(m, m, visc=0.01, rhof=1.0, tw=2.5, pinw=2.5)
    for cx, cy, r in sediment_circles(count, seed, m):
        f.add_circle(cx, cy, r, 1.0, G=1.0, rhos=2.0, ev_trans_mult=1.0)
    f.set_solver(rel_tol=rel_tol, max_cycles=max_cycles, precision=precision)
    f.initialize(0.8, 0.4, 0.8)
    return f


# ------------------------------------------------------------ learned guess
def _f(a):
    return a.ctypes.data_as(_fp)


class UNet:
    """The learned projection guess network (include/incrmt.h): a 4-level
    U-Net in BF16 on the MFMA units for an nx x ny node grid. Parameters are
    one packed float32 vector (per layer weights [cout][tap][cin], bias)."""

    def __init__(self, nx, ny, seed=1):
        h = _V()
        _check(lib().incrmt_unet_create(nx, ny, seed, ctypes.byref(h)))
        self._h, self.nx, self.ny = h, nx, ny

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            lib().incrmt_unet_destroy(self._h)
            self._h = None

    @property
    def nparams(self):
        n = ctypes.c_longlong(0)
        _check(lib().incrmt_unet_nparams(self._h, ctypes.byref(n)))
        return n.value

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().incrmt_unet_get_params(self._h, _f(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        _check(lib().incrmt_unet_set_params(self._h, _f(p)))

    def _in(self, in3):
        a = np.ascontiguousarray(in3, dtype=np.float32)
        assert a.size == self.nx * self.ny * 3
        return a

    def forward(self, in3):
        """in3: (ny, nx, 3) float32; returns the (ny, nx) prediction."""
        a = self._in(in3)
        out = np.zeros(self.nx * self.ny, dtype=np.float32)
        _check(lib().incrmt_unet_forward(self._h, _f(a), _f(out)))
        return out.reshape(self.ny, self.nx)

    def gradient(self, in3, target):
        a = self._in(in3)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = _D(0)
        g = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().incrmt_unet_gradient(self._h, _f(a), _f(t), ctypes.byref(loss), _f(g)))
        return loss.value, g

    def train_step(self, in3, target, lr):
        a = self._in(in3)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = _D(0)
        _check(lib().incrmt_unet_train_step(self._h, _f(a), _f(t), lr, ctypes.byref(loss)))
        return loss.value


def guess_policy_run(rel, tol=1e-4):
    """The learned start's accept/fallback rule (host logic, no GPU) over a
    sequence of relative residuals after the first V-cycle: 1 accepted, 0
    fallback, -1 learned start switched off."""
    r = np.ascontiguousarray(rel, dtype=np.float64)
    d = np.zeros(r.size, dtype=np.int32)
    _check(lib().incrmt_guess_policy_run(_d(r), r.size, tol, d.ctypes.data_as(_ip)))
    return d
