"""Discrete divergences of the step's projected velocities, from the
reference's own source-term formulas applied to the projected fields.

FEM (fluid_2d.cc:1297-1309): at node (i,j) of the (m+1)x(n+1) corner grid,
  hx (c0[i-1,j-1] + c0[i-1,j] - c0[i,j-1] - c0[i,j])
+ hy (c1[i-1,j-1] - c1[i-1,j] + c1[i,j-1] - c1[i,j]),
with cells outside the box counting zero (no-slip, fluid_2d.cc:839-872).
MAC (fluid_2d.cc:1179-1187): xsp (ul[i+1,j]-ul[i,j]) + ysp (vd[i,j+1]-vd[i,j]).
Arrays are on the ghosted (m+4)x(n+4) grid.
"""
import numpy as np


def fem_div(c0, c1, m, n, dx, dy, dt=1.0):
    a = np.zeros((n + 2, m + 2))
    b = np.zeros((n + 2, m + 2))
    a[1:-1, 1:-1] = c0[2:-2, 2:-2]
    b[1:-1, 1:-1] = c1[2:-2, 2:-2]
    hx, hy = 0.5 * dy / dt, 0.5 * dx / dt
    # node (i,j) touches padded cells (i,j),(i,j+1),(i+1,j),(i+1,j+1)
    dl, ul_, dr, ur = a[:-1, :-1], a[1:, :-1], a[:-1, 1:], a[1:, 1:]
    ed, eu, edr, eur = b[:-1, :-1], b[1:, :-1], b[:-1, 1:], b[1:, 1:]
    return hx * (dl + ul_ - dr - ur) + hy * (ed - eu + edr - eur)


def mac_div(ul, vd, m, n, dx, dy):
    u = ul[2:-2, 2:m + 3]
    v = vd[2:n + 3, 2:-2]
    return (u[:, 1:] - u[:, :-1]) / dx + (v[1:, :] - v[:-1, :]) / dy
