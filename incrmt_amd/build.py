"""Build the HIP library libincrmt_hip.so in-tree for gfx950.

The library is compiled without floating-point contraction so that the
stencil sweeps reproduce the reference's FP64 arithmetic operation by
operation, and without packed FP32 instructions.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libincrmt_hip.so")
SOURCES = ["step.hip", "mg.hip", "levelset.hip", "tracers.hip", "unet.hip", "engine.hip"]
ARCH = os.environ.get("INCRMT_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src, objdir):
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(lambda s: _compile(s, objdir), SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for (_, err), src in zip(results, SOURCES):
            warn = [l for l in err.splitlines() if "warning" in l and "packed-fp32" not in l]
            if warn:
                print(src, "\n".join(warn[:20]), file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-z,defs", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_driver()
    return LIB


DRIVER = os.path.join(HERE, "adapter_driver")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_driver(out=DRIVER):
    """The C++ driver written against include/incrmt_adapter.hh (the
    reference-side adapters), linked to the library beside it."""
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++11", "-Wall", "-I", INCLUDE, os.path.join(CSRC, "adapter_driver.cc"),
           "-L", HERE, "-lincrmt_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"adapter driver build failed\n{r.stdout}\n{r.stderr}")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
