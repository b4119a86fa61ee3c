"""Frame output (fluid_2d::write_files, fluid_2d_io.cc:60-179): the GPU
engine's files against the same fields of the oracle written out by the
reference's rules, and a closed form (a rigid rotation's vorticity is 2w).
Stated tolerance: headers (grid coordinates) identical; values within 1e-5
relative to the field's scale (float32 storage plus the per-step parity
tolerance)."""
import math
import os

import numpy as np
import pytest

import incrmt_amd as ia
from tests import frames as fr
from tests import oracle_lib as ol
from tests.test_gpu_step import gpu_from_spec

pytestmark = pytest.mark.gpu


def test_frames_match_oracle_fields(tmp_path):
    m = 64
    spec = ol.simple_spin_spec(m)
    g, o = gpu_from_spec(spec), ol.OracleFluid(spec, mg_tol=1e-13)
    l, adt = g.timestep_select(4 * math.pi / 240)
    for _ in range(6):
        g.step_forward(adt)
        o.step(adt)
    g.write_frame(tmp_path, 3, 1 | 2 | 8 | 16 | 64 | 128 | 256 | 512 | 4096)
    h = 2.0 / m
    ml = m + 4
    u, v, p = o.get("u"), o.get("v"), o.get("p")
    # cell-centred fields on the interior cells, corner fields on m+1 corners
    x, y, U = fr.read_matrix(os.path.join(tmp_path, "u.3"))
    assert np.array_equal(x, np.float32(-1 + h * (np.arange(m) + 0.5)))
    assert np.array_equal(y, np.float32(-1 + h * (np.arange(m) + 0.5)))
    assert np.abs(U - u[2:-2, 2:-2]).max() <= 1e-5 * np.abs(u).max()
    x, y, P = fr.read_matrix(os.path.join(tmp_path, "p.3"))
    assert P.shape == (m + 1, m + 1)
    assert np.array_equal(x, np.float32(-1 + h * np.arange(m + 1)))
    assert np.abs(P - p[2:-1, 2:-1]).max() <= 1e-5 * np.abs(p).max()
    w = fr.corner_vorticity(u, v, 1 / h, 1 / h)
    _, _, W = fr.read_matrix(os.path.join(tmp_path, "w.3"))
    assert np.abs(W - w[2:-1, 2:-1]).max() <= 1e-5 * np.abs(w).max()
    # solid fields on the band's bounding box, padded by 2
    s = o.get("s")
    J, I = np.nonzero(s < 4)
    li, ui, lj, uj = max(I.min() - 2, 0), min(I.max() + 3, ml), max(J.min() - 2, 0), min(J.max() + 3, ml)
    x, y, PHI = fr.read_matrix(os.path.join(tmp_path, "phi.3"))
    assert PHI.shape == (uj - lj, ui - li)
    assert np.array_equal(x, np.float32(-1 + h * (np.arange(li, ui) - 1.5)))
    assert np.abs(PHI - o.get("phi")[lj:uj, li:ui]).max() <= 0.05 * h
    _, _, X = fr.read_matrix(os.path.join(tmp_path, "X.3"))
    assert np.abs(X - o.get("X")[lj:uj, li:ui]).max() <= 1e-6
    # div is evaluated at lower-left corners but written over the m x m
    # interior with cell-centre coordinates: the reference's output() keys
    # its region and coordinates on the mode bits 3, 4, 9, 11 only
    # (fluid_2d_io.cc:113, 130)
    x, _, D = fr.read_matrix(os.path.join(tmp_path, "div.3"))
    assert D.shape == (m, m)
    assert np.array_equal(x, np.float32(-1 + h * (np.arange(m) + 0.5)))


def test_rigid_rotation_vorticity(tmp_path):
    m = 32
    g = ia.simple_spin(m, rel_tol=1e-10)
    h = 2.0 / m
    c = -1 + h * (np.arange(m + 4) - 1.5)
    X, Y = np.meshgrid(c, c)
    om = 0.7
    g.set("u", -om * Y)
    g.set("v", om * X)
    g.write_frame(tmp_path, 0, 16 | 4096)
    _, _, W = fr.read_matrix(os.path.join(tmp_path, "w.0"))
    assert np.allclose(W, 2 * om, rtol=1e-6)
    _, _, D = fr.read_matrix(os.path.join(tmp_path, "div.0"))
    assert D.shape == (m, m) and np.abs(D).max() <= 1e-6


def test_solve_frame_lines(tmp_path, capfd):
    """fluid_2d::solve's frame loop (fluid_2d.cc:259-296): the header file,
    frame 0, then per frame l steps, the frame's files and one line: with
    fflags & 8192 "frame steps t_step t_out  mac_vcycles mac_s fem_vcycles
    fem_s" (the V-cycles of the frame, which the step statistics count too),
    else "# Output frame k [l, ...] {MAC avg, FEM avg}"."""
    f = ia.simple_spin(64, rel_tol=1e-8)
    l, adt = f.timestep_select(0.02)
    c0 = f.stats()
    f.solve(0.04, 2, tmp_path, fflags=8 | 8192)
    c1 = f.stats()
    f.solve(0.02, 1, tmp_path, fflags=8)
    out = capfd.readouterr().out.strip().splitlines()
    assert out[0] == "# Output frame 0"
    rows = [line.split() for line in out[1:3]]
    assert [int(r[0]) for r in rows] == [1, 2] and all(int(r[1]) == l for r in rows)
    assert all(len(r) == 8 and float(r[2]) > 0 and float(r[5]) > 0 and float(r[7]) > 0 for r in rows)
    assert sum(int(r[4]) for r in rows) == c1["total_mac_cycles"] - c0["total_mac_cycles"]
    assert sum(int(r[6]) for r in rows) == c1["total_fem_cycles"] - c0["total_fem_cycles"]
    assert out[3].startswith(f"# Output frame 3 [{l}, ") and "{MAC " in out[3] and ", FEM " in out[3]
    hdr = (tmp_path / "header").read_text().split("\n")
    t0, t1, fr = hdr[0].split()
    assert float(t0) == 0 and abs(float(t1) - 0.04) < 1e-12 and fr == "2"
    assert hdr[1].split()[2] == "1"
    for k in range(4):
        assert (tmp_path / f"p.{k}").exists()
    f.close()
