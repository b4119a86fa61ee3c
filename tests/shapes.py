"""Level-set test shapes on the reference's ghosted grid geometry.

A fluid grid of m x n cells on [ax,bx]x[ay,by] carries level sets on an
(m+4) x (n+4) grid whose point (I,J) sits at the cell centre of cell
(I-2, J-2) (obj_field.cc:18, fluid_2d.cc:203-206).
"""
import math

import numpy as np


def ls_geometry(m, n, ax=-1.0, bx=1.0, ay=-1.0, by=1.0):
    dx = (bx - ax) / m
    dy = (by - ay) / n
    return dict(m=m + 4, n=n + 4, ax=ax - 1.5 * dx, ay=ay - 1.5 * dy, dx=dx, dy=dy)


def centres(geo):
    x = geo["ax"] + geo["dx"] * np.arange(geo["m"])
    y = geo["ay"] + geo["dy"] * np.arange(geo["n"])
    return np.meshgrid(x, y)  # X[j,i], Y[j,i]


def rotor_consts(r_scale=0.8, a1=0.5, a2=0.125):
    scf = r_scale / (1 + a1 + a2)
    return scf * (1 - a2), scf * a1, 2 * scf * a2


def rotor_phi(X, Y, r_scale=0.8, a1=0.5, a2=0.125):
    """obj_simple_spin::phi (object.cc:167-171) evaluated with numpy."""
    c0, c1, c2 = rotor_consts(r_scale, a1, a2)
    th = np.arctan2(Y, X)
    c = np.cos(3 * th)
    return np.sqrt(X * X + Y * Y) - c0 - c * (c1 + c2 * c)


def rotated(X, Y, ang):
    ca, sa = math.cos(ang), math.sin(ang)
    return ca * X + sa * Y, -sa * X + ca * Y


def shape(kind, geo, ang=0.0):
    X, Y = centres(geo)
    if kind == "rotor":
        Xr, Yr = rotated(X, Y, ang)
        return rotor_phi(Xr, Yr)
    if kind == "circle":
        return np.sqrt((X - 0.1) ** 2 + (Y + 0.05) ** 2) - 0.5
    if kind == "wavy":
        return np.sqrt((X - 0.08) ** 2 + Y * Y) - 0.4 + 0.18 * np.sin(11 * X) + 0.13 * np.sin(13.5 * Y)
    if kind == "two":
        a = np.sqrt((X - 0.35) ** 2 + (Y - 0.1) ** 2) - 0.25
        b = np.sqrt((X + 0.35) ** 2 + (Y + 0.1) ** 2) - 0.3
        return np.minimum(a, b)
    raise ValueError(kind)
