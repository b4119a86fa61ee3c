"""The drop-in boundary: the C-ABI library loads and exports every entry
point include/incrmt.h declares (no compute calls: these run without a GPU),
and the product never reaches the oracle."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "incrmt.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(incrmt_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_interface():
    names = declared()
    for must in ("incrmt_fluid_step", "incrmt_mg_solve", "incrmt_levelset_build_band",
                 "incrmt_levelset_extrapolate", "incrmt_fluid_create", "incrmt_fluid_download"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import incrmt_amd as ia
    lib = ia.lib()
    for name in declared():
        assert hasattr(lib, name), name
    assert sorted(ia.SYMBOLS) == declared()


def test_library_is_built_for_gfx950():
    # the HIP fat binary of this repository's own build names its target
    import incrmt_amd as ia
    blob = open(ia.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_product_does_not_use_the_oracle():
    for dirpath, _, files in os.walk(os.path.join(ROOT, "incrmt_amd")):
        for fn in files:
            if fn.endswith((".py", ".hip", ".hh", ".cc", ".h")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in text.lower(), fn
    import incrmt_amd as ia
    ldd = subprocess.run(["ldd", ia.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in ldd and "reflevel" not in ldd


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import incrmt_amd as ia
    monkeypatch.setattr(ia, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(ia, "_lib", None)
    with pytest.raises(ia.IncrmtError):
        ia.lib()


def test_status_codes_match_header():
    import incrmt_amd as ia
    src = open(os.path.join(ROOT, "include", "incrmt.h")).read()
    codes = dict(re.findall(r"#define (INCRMT_ERR_[A-Z_]+) (\d+)", src))
    assert int(codes["INCRMT_ERR_NOT_CONVERGED"]) == ia.ERR_NOT_CONVERGED
    assert int(codes["INCRMT_ERR_NEWTON"]) == ia.ERR_NEWTON
    assert int(codes["INCRMT_ERR_SINGULAR_FIT"]) == ia.ERR_SINGULAR_FIT
    assert int(codes["INCRMT_ERR_HIP"]) == ia.ERR_HIP


def test_library_resolves_its_own_symbols():
    # every C++ symbol of the engine is defined inside the library (a missing
    # definition would only surface when the loader binds it)
    import incrmt_amd as ia
    r = subprocess.run(["nm", "-D", "--undefined-only", "-C", ia.LIB_PATH], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "incrmt::" not in r.stdout
