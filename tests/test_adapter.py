"""The reference-side C++ adapters (include/incrmt_adapter.hh: fluid_2d's
set-up and step loop, mgs_base's solver surface over the C ABI) compile as
C++11 against include/incrmt.h and link against the library with the
driver written on them (incrmt_amd/csrc/adapter_driver.cc); without a GPU
the driver is only built and its usage checked (tests/test_gpu_adapter.py
runs it)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "incrmt_amd", "libincrmt_hip.so")


def test_adapter_header_is_cxx11():
    r = subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only", "-x", "c++", "-I",
                        os.path.join(ROOT, "include"), os.path.join(ROOT, "include", "incrmt_adapter.hh")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_adapter_driver_builds_and_links(tmp_path):
    from incrmt_amd import build as hb
    out = hb.build_driver(str(tmp_path / "adapter_driver"))
    # the library (and its HIP runtime) resolve; run without arguments: usage
    shutil.copy(LIB, tmp_path / "libincrmt_hip.so")
    r = subprocess.run([out], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
