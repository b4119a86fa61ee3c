"""The closed-form pins of tests/test_oracle_step_pins.py catch plausible
slips in the oracle's step sweeps.

Each case plants one mistake in a copy of oracle/fluid_oracle.cc (a swapped
term, a wrong sign, a dropped factor, the wrong side of an upwind choice),
builds that copy into its own library in a temporary folder, and runs the
pin tests (and the invariants of tests/test_oracle_step.py) against it:
every planted mistake must make them fail. The
unmodified oracle passes the same tests (they run in the ordinary suite).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
FLAGS = ["-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c++17", "-w"]

# (name, reference line, text in fluid_oracle.cc, planted text, occurrence)
MUTANTS = [
    ("half-time edge: c0 for c2", "fluid_2d.cc:992",
     "f.ul = f.u + 0.5 * (-dx * f.ux + dt * (-f.u * f.ux - f.c2 + accx));",
     "f.ul = f.u + 0.5 * (-dx * f.ux + dt * (-f.u * f.ux - f.c0 + accx));", 0),
    ("half-time edge: dt/2 on the advective term", "fluid_2d.cc:1000",
     "f.ud = f.u + 0.5 * (-dy * f.uy + dt * (-f.v * f.uy - f.c0 + accx));",
     "f.ud = f.u + 0.5 * (-dy * f.uy + 0.5 * dt * (-f.v * f.uy - f.c0 + accx));", 0),
    ("stress: dropped +1 in the trace term", "fluid_2d.cc:544",
     "double sfac = (1 / 3.0) * ((sa + sd) + 1)", "double sfac = (1 / 3.0) * ((sa + sd))", 0),
    ("stress: F component swap", "fluid_2d.cc:539",
     "double Fa = Yy, Fb = -Xy, Fc = -Yx, Fd = Xx;", "double Fa = Yy, Fb = -Yx, Fc = -Xy, Fd = Xx;", 0),
    ("fluid stress: sign in the shear term", "fluid_2d.cc:589",
     "s2f = visc * (vx + uy);", "s2f = visc * (vx - uy);", 0),
    ("solid extra viscosity: y spacing on an x difference", "fluid_2d.cc:552",
     "s1s += ex_visc * xsp * (fp->u - fp[-1].u);", "s1s += ex_visc * ysp * (fp->u - fp[-1].u);", 0),
    ("u*: sign of dt acc", "fluid_2d.cc:1287",
     "- hy * (fu.vd + f.vd) * (fu.ud - f.ud) + dt * accx;", "- hy * (fu.vd + f.vd) * (fu.ud - f.ud) - dt * accx;", 0),
    ("u*: hx for hy", "fluid_2d.cc:1288",
     "f.c1 = f.v - hx * (fr.ul + f.ul) * (fr.vl - f.vl) - hy * (fu.vd + f.vd) * (fu.vd - f.vd)",
     "f.c1 = f.v - hx * (fr.ul + f.ul) * (fr.vl - f.vl) - hx * (fu.vd + f.vd) * (fu.vd - f.vd)", 0),
    ("acceleration: no division by rho", "fluid_2d.cc:1045",
     "accx /= f.rho; accy /= f.rho;", "accy /= f.rho;", 0),
    ("FEM source: dx and dy exchanged", "fluid_2d.cc:1301",
     "double hx = 0.5 * dy / dt, hy = 0.5 * dx / dt;", "double hx = 0.5 * dx / dt, hy = 0.5 * dy / dt;", 0),
    ("pressure update: v's gradient in u", "fluid_2d.cc:1362",
     "fp->u = u_star - 0.5 * dt * rhoinv * xsp * (fp[ml + 1].p + fp[1].p - fp[ml].p - fp->p);",
     "fp->u = u_star - 0.5 * dt * rhoinv * xsp * (fp[ml + 1].p - fp[1].p + fp[ml].p - fp->p);", 0),
    ("copy_pressure: no half weight on the wall rows", "fluid_2d.cc:1344",
     "if (!sp.y_prd && (j == 0 || j == n)) prow *= 0.5;", "if (!sp.y_prd && (j == 0 || j == n)) prow *= 1.0;", 0),
    ("limiter: 3 for 4 in the fourth-order slope", "fluid_2d.cc:606",
     "double s = (1 / 6.) * (4 * tdrc", "double s = (1 / 6.) * (3 * tdrc", 0),
    ("edge extrapolant: sign of the dt u term", "fluid_2d.cc:895",
     "f.vl = f.v - 0.5 * (dx + dt * f.u) * f.vx;", "f.vl = f.v - 0.5 * (dx - dt * f.u) * f.vx;", 0),
    ("tangential term: reversed difference", "fluid_2d.cc:961",
     "f.c1 = u_adv * (fr.vl - f.vl);", "f.c1 = u_adv * (f.vl - fr.vl);", 0),
    ("upwinding: keeps the downwind tangential value", "fluid_2d.cc:672",
     "fu.vd = f.vu; fu.ud = f.uu;", "fu.vd = f.vu; fu.ud = fu.ud;", 0),
    ("MAC correction: wrong sign on x faces", "fluid_2d.cc:1207",
     "if (i > 0) f.ul -= (*so - so[-1]) * xsp;", "if (i > 0) f.ul += (*so - so[-1]) * xsp;", 0),
    ("MAC correction, varying density: wrong face density", "fluid_2d.cc:1227",
     "if (i > 0) f.ul -= el * (*so - so[-1]) * xsp;", "if (i > 0) f.ul -= ed * (*so - so[-1]) * xsp;", 0),
    ("reference map: Yy for Xy in the half-time X", "fluid_2d.cc:1092",
     "double Xdt = dt * (-f.u * s.Xx - f.v * s.Xy)", "double Xdt = dt * (-f.u * s.Xx - f.v * s.Yy)", 0),
    ("reference map update: missing factor 1/2 in hy", "fluid_2d.cc:1246",
     "double hx = 0.5 * dt * xsp, hy = 0.5 * dt * ysp;\n    for (auto &o : obj) {",
     "double hx = 0.5 * dt * xsp, hy = dt * ysp;\n    for (auto &o : obj) {", 0),
]


def _plant(src, old, new, occ):
    k = -1
    for _ in range(occ + 1):
        k = src.index(old, k + 1)
    return src[:k] + new + src[k + len(old):]


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    common = []
    for f in ("oracle.cc", "levelset_oracle.cc", "mg_oracle.cc"):
        o = str(d / (f + ".o"))
        common.append((["g++"] + FLAGS + ["-I", ORACLE, "-c", os.path.join(ORACLE, f), "-o", o], o))
    src = open(os.path.join(ORACLE, "fluid_oracle.cc")).read()
    jobs = []
    for k, (name, cite, old, new, occ) in enumerate(MUTANTS):
        assert src.count(old) >= occ + 1, f"mutant '{name}': text not found in fluid_oracle.cc"
        path = d / f"fluid_mut{k}.cc"
        path.write_text(_plant(src, old, new, occ))
        jobs.append((["g++"] + FLAGS + ["-I", ORACLE, "-c", str(path), "-o", str(path) + ".o"], str(path) + ".o"))

    def run(cmd):
        subprocess.run(cmd[0], check=True, capture_output=True)
        return cmd[1]

    with ThreadPoolExecutor(max_workers=6) as ex:
        objs = list(ex.map(run, common + jobs))
    base = objs[:3]
    libs = []
    for k, o in enumerate(objs[3:]):
        lib = str(d / f"liboracle_mut{k}.so")
        subprocess.run(["g++", "-shared", "-fopenmp", "-o", lib] + base + [o], check=True, capture_output=True)
        libs.append(lib)
    return libs


def test_every_planted_slip_fails_a_pin(mutant_libs):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    pins = [os.path.join(ROOT, "tests", t) for t in ("test_oracle_step_pins.py", "test_oracle_step.py")]

    def run(k):
        e = dict(env, INCRMT_ORACLE_SO=mutant_libs[k])
        r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] + pins,
                           cwd=ROOT, env=e, capture_output=True, text=True, timeout=600)
        return r.returncode

    with ThreadPoolExecutor(max_workers=6) as ex:
        rcs = list(ex.map(run, range(len(MUTANTS))))
    survived = [MUTANTS[k][0] for k, rc in enumerate(rcs) if rc == 0]
    assert not survived, f"planted slips no pin catches: {survived}"
    assert all(rc == 1 for rc in rcs), rcs   # failed tests, not crashes of the harness
