"""The seeded workload generators (incrmt_amd/workloads.py)."""
import numpy as np

from incrmt_amd import workloads as wl
from tests import operators as ops


def test_mt19937_64_matches_the_standard():
    # C++11 [rand.predef]: the 10000th output of a default-constructed
    # std::mt19937_64 is 9981545732273789042
    r = wl.MT19937_64()
    for _ in range(9999):
        r.next()
    assert r.next() == 9981545732273789042


def test_uniform_and_normal_draws():
    r = wl.MT19937_64(1)
    u = np.array([r.uniform() for _ in range(20000)])
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.01
    g = np.array([r.normal() for _ in range(20000)])
    assert abs(g.mean()) < 0.03 and abs(g.std() - 1) < 0.03


def test_fem_source_is_the_operator_image_of_a_discrete_gradient():
    # closed form: for a linear nodal field p the cell-centred gradient is its
    # exact gradient, so the source of u = grad p is A p with the
    # bilinear-FEM stiffness (both integrate grad p . grad w over the box)
    m, n = 9, 7
    h = 0.125
    yy, xx = np.mgrid[0:n + 1, 0:m + 1] * h
    p = 0.3 + 1.7 * xx - 0.6 * yy
    u = 0.5 * (p[:-1, 1:] + p[1:, 1:] - p[:-1, :-1] - p[1:, :-1]) / h
    v = 0.5 * (p[1:, :-1] + p[1:, 1:] - p[:-1, :-1] - p[:-1, 1:]) / h
    b = wl.fem_source(u, v, h, h)
    A = ops.fem_matrix(m + 1, n + 1, h, h)
    assert np.allclose(b.ravel(), A @ p.ravel(), atol=1e-12)


def test_fourier_rhs_is_compatible():
    b, h = wl.fourier_fem_rhs(65)
    assert b.shape == (65, 65) and h == 1 / 64
    assert abs(b.sum()) <= 1e-14 * np.abs(b).sum()
