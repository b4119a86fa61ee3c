"""Synthetic inputs of the benchmark configurations, generated from seeds.

Every random draw comes from the 64-bit Mersenne Twister (std::mt19937_64,
restated here so the inputs do not depend on a C++ library's distribution
code), mapped to uniforms as (x >> 11) / 2^53 and to normals by the
Box-Muller transform; the exact recipes are spelled out per function.
"""
import math

import numpy as np

_MASK = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne Twister of C++11 std::mt19937_64 (w=64, n=312,
    m=156, r=31); seeded as the standard specifies."""

    def __init__(self, seed=5489):
        self.mt = [0] * 312
        self.mt[0] = seed & _MASK
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & _MASK
        self.idx = 312

    def _twist(self):
        mt = self.mt
        for i in range(312):
            x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self.idx = 0

    def next(self):
        if self.idx >= 312:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & _MASK

    def uniform(self):
        """A double in [0, 1): the top 53 bits."""
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self):
        """Box-Muller on two uniforms: sqrt(-2 ln(1-u1)) cos(2 pi u2)."""
        u1, u2 = self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)

    def below(self, n):
        """An integer in [0, n): next() mod n (the bias is below 2^-50 here)."""
        return self.next() % n


def fourier_velocity(nodes, seed=20260101, modes=32, kmax=16):
    """Cell-centred velocity of the standalone FEM configuration on the unit
    square with nodes x nodes grid nodes (h = 1/(nodes-1)): u, v are sums of
    `modes` Fourier modes sin(pi k1 x) sin(pi k2 y), with (k1,k2) drawn from
    {1..kmax}^2 without replacement (a partial Fisher-Yates shuffle of the
    row-major list, index j swapped with j + below(N-j)), and amplitudes
    a_k, b_k = normal() / |k|^2 drawn in the order a_1, b_1, a_2, b_2, ...
    Returns (u, v) of shape (nodes-1, nodes-1), row j = y index."""
    rng = MT19937_64(seed)
    pool = [(k1, k2) for k2 in range(1, kmax + 1) for k1 in range(1, kmax + 1)]
    n = len(pool)
    for j in range(modes):
        r = j + rng.below(n - j)
        pool[j], pool[r] = pool[r], pool[j]
    ks = pool[:modes]
    amps = []
    for k1, k2 in ks:
        kk = k1 * k1 + k2 * k2
        amps.append((rng.normal() / kk, rng.normal() / kk))
    m = nodes - 1
    h = 1.0 / m
    xc = (np.arange(m) + 0.5) * h
    u = np.zeros((m, m))
    v = np.zeros((m, m))
    for (k1, k2), (a, b) in zip(ks, amps):
        sx = np.sin(math.pi * k1 * xc)[None, :]
        sy = np.sin(math.pi * k2 * xc)[:, None]
        u += a * (sy * sx)
        v += b * (sy * sx)
    return u, v, h


def fem_source(u, v, dx, dy, dt=1.0):
    """The nodal right-hand side of finite_element_projection
    (fluid_2d.cc:1297-1309) for cell values c0 = u, c1 = v with the walls'
    zero ghost cells (fluid_2d.cc:839-872): node (i,j) sees cells
    (i-1,j-1), (i-1,j), (i,j-1), (i,j). Returns an (m+1) x (n+1) array."""
    n, m = u.shape
    c0 = np.zeros((n + 2, m + 2))
    c1 = np.zeros((n + 2, m + 2))
    c0[1:-1, 1:-1] = u
    c1[1:-1, 1:-1] = v
    hx, hy = 0.5 * dy / dt, 0.5 * dx / dt
    dl0, ul0, dr0, ur0 = c0[:-1, :-1], c0[1:, :-1], c0[:-1, 1:], c0[1:, 1:]
    dl1, ul1, dr1, ur1 = c1[:-1, :-1], c1[1:, :-1], c1[:-1, 1:], c1[1:, 1:]
    return hx * (dl0 + ul0 - dr0 - ur0) + hy * (dl1 - ul1 + dr1 - ur1)


def fourier_fem_rhs(nodes=1024, seed=20260101):
    """Right-hand side of the standalone FEM configuration: the nodal
    divergence (as fem_source forms it, dt = 1) of fourier_velocity."""
    u, v, h = fourier_velocity(nodes, seed)
    return fem_source(u, v, h, h), h


def sediment_circles(count=64, seed=20260102, m=2048, ax=-1.0, bx=1.0, ay=-1.0, by=1.0,
                     rmin=0.02, rmax=0.05):
    """The multibody configuration's solids: `count` circles with radii
    uniform in [rmin, rmax] times the domain width, centres uniform in the
    box such that every circle keeps a gap of 2.5 sqrt(2) h from the walls
    and from every other circle (the sediment.cc:82 padding rule, h the
    cell width of an m x m grid). Draws per attempt: radius, x, y; an
    attempt that violates a gap is discarded (its radius too).
    Returns a list of (cx, cy, r)."""
    rng = MT19937_64(seed)
    w = bx - ax
    gap = 2.5 * math.sqrt(2.0) * (w / m)
    out = []
    tries = 0
    while len(out) < count:
        tries += 1
        if tries > 1000000:
            raise RuntimeError("cannot place the circles")
        r = (rmin + (rmax - rmin) * rng.uniform()) * w
        lo_x, hi_x = ax + r + gap, bx - r - gap
        lo_y, hi_y = ay + r + gap, by - r - gap
        cx = lo_x + (hi_x - lo_x) * rng.uniform()
        cy = lo_y + (hi_y - lo_y) * rng.uniform()
        if all(math.hypot(cx - x, cy - y) >= r + s + gap for x, y, s in out):
            out.append((cx, cy, r))
    return out


def rotor_ensemble(count=256, seed=20260103, sizes=(64, 128, 256, 512)):
    """The batched-ensemble configuration: per instance a grid size drawn
    from `sizes` (below(len(sizes))) and then a rotor angular velocity
    uniform in [0.5, 2]. Returns a list of (m, omega)."""
    rng = MT19937_64(seed)
    out = []
    for _ in range(count):
        m = sizes[rng.below(len(sizes))]
        om = 0.5 + 1.5 * rng.uniform()
        out.append((m, om))
    return out
