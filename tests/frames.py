"""Reader of the reference's frame files (fluid_2d::output, fluid_2d_io.cc:
119-179): float32 Gnuplot matrices whose first row holds the column count
and the x coordinates and whose other rows hold y and the values. Test
infrastructure."""
import numpy as np


def read_matrix(path):
    a = np.fromfile(path, dtype=np.float32)
    si = int(a[0])
    rows = a.reshape(-1, si + 1)
    x = rows[0, 1:]
    y = rows[1:, 0]
    return x, y, rows[1:, 1:]


def corner_vorticity(u, v, xsp, ysp):
    """vorticity() at the lower-left corner of every cell (fluid_2d.hh:252-255),
    arrays on the ghosted grid; entry [J, I] for J, I >= 1."""
    w = np.zeros_like(u)
    w[1:, 1:] = 0.5 * (xsp * (v[1:, 1:] - v[1:, :-1] + v[:-1, 1:] - v[:-1, :-1])
                       - ysp * (u[1:, 1:] + u[1:, :-1] - u[:-1, 1:] - u[:-1, :-1]))
    return w
