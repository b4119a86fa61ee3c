"""The GPU band rebuild evaluates the eikonal local update in a
branch-reduced form and runs the inward march as the negated outward one
(incrmt_amd/csrc/eikonal.hh, levelset.hip). Both must equal the
reference-shaped update (heap.cc:177-388) bit for bit; this compiles the
product's header on the host and checks that on a million random
neighbourhoods."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_branch_reduced_and_sign_flipped_updates_are_bit_identical(tmp_path):
    exe = str(tmp_path / "eikonal_forms")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", os.path.join(HERE, "eikonal_forms.cc"),
                    "-o", exe], check=True)
    r = subprocess.run([exe, "1000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split() == ["0", "0"]
