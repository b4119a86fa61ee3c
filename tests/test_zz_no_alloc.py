"""incrmt_levelset_build_band_noalloc, the band rebuild that allocates no
device memory, against incrmt_levelset_build_band (which allocates its own
work buffers) through the C ABI: identical phi, status codes and pass
counts, on two calls in a row (the second on a smaller grid, so the shared
scratch array still holds the first call's larger contents), and a clear
error beyond the size limit. The caller's arrays are device memory that
torch allocates; the entry point itself allocates none."""
import ctypes

import numpy as np
import pytest
import torch

import incrmt_amd as ia
from tests import shapes

pytestmark = pytest.mark.gpu


def noalloc(args, phi):
    m, n = args[0], args[1]
    dphi = torch.from_numpy(np.ascontiguousarray(phi, dtype=np.float64)).to("cuda:0")
    ds = torch.full((m * n,), -1, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    passes = ctypes.c_int(0)
    rc = ia.lib().incrmt_levelset_build_band_noalloc(*args, ctypes.c_void_p(dphi.data_ptr()),
                                                     ctypes.c_void_p(ds.data_ptr()), ctypes.byref(passes))
    return rc, dphi.cpu().numpy(), ds.cpu().numpy(), passes.value


@pytest.mark.parametrize("sizes", [((512, 512), (200, 150)), ((160, 96), (64, 64))])
def test_noalloc_matches_allocating_entry_point(sizes):
    for k, (m, n) in enumerate(sizes):
        geo = shapes.ls_geometry(m, n)
        lim = 7.0 * geo["dx"]
        args = (geo["m"], geo["n"], geo["ax"], geo["ay"], geo["dx"], geo["dy"], -lim, lim)
        phi = shapes.shape("rotor" if k == 0 else "wavy", geo, 0.7).ravel()
        print(f"no-alloc call {k}: level-set grid {geo['m']} x {geo['n']}", flush=True)
        pa, sa, na = ia.build_band(*args, phi)
        rc, pb, sb, nb = noalloc(args, phi)
        assert rc == 0, ia.lib().incrmt_last_error().decode()
        assert np.array_equal(sa, sb)
        assert np.array_equal(pa, pb)
        assert na == nb


def test_noalloc_rejects_oversize_and_host_arrays():
    geo = shapes.ls_geometry(600, 64)
    lim = 7.0 * geo["dx"]
    args = (geo["m"], geo["n"], geo["ax"], geo["ay"], geo["dx"], geo["dy"], -lim, lim)
    phi = shapes.shape("circle", geo, 0.0).ravel()
    rc, _, _, _ = noalloc(args, phi)
    assert rc == ia.ERR_ARG
    assert b"no-allocation limit" in ia.lib().incrmt_last_error()
    geo = shapes.ls_geometry(64, 64)
    args = (geo["m"], geo["n"], geo["ax"], geo["ay"], geo["dx"], geo["dy"], -lim, lim)
    hphi = np.zeros(geo["m"] * geo["n"])
    hs = np.zeros(geo["m"] * geo["n"], dtype=np.int32)
    rc = ia.lib().incrmt_levelset_build_band_noalloc(*args, hphi.ctypes.data_as(ctypes.c_void_p),
                                                     hs.ctypes.data_as(ctypes.c_void_p), None)
    assert rc == ia.ERR_ARG
