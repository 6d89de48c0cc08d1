"""The four cases of spatialops/structured/test/testNeboBC.cpp as data: field
layouts, the fill of phi and gamma, the minus/plus mask points and the expected
phi after both conditions, each computed with the test's own formula.

Shared by the CPU (oracle) and GPU tests of the boundary-condition builder.
"""
from __future__ import annotations

import numpy as np

import nebo_hip as nh

DIM = (3, 3, 1)
LENGTH = 1.5
DX = LENGTH / DIM[0]
DY = LENGTH / DIM[1]
BC = (True, True, True)


def _grid(layout, fn):
    """Values fn(i, j) over the window, i and j counted from -1 like the test's loops."""
    g = layout.glob
    a = np.empty((g[2], g[1], g[0]))
    for j in range(g[1]):
        for i in range(g[0]):
            a[:, j, i] = fn(i - 1, j - 1)
    return a


def case(name):
    """(operator key, phi layout, gamma layout, phi, gamma, expected phi,
    minus points, plus points) of one test block."""
    db = nh.build_stencils(*DIM, LENGTH, LENGTH, LENGTH)
    if name == "SVol GradX":          # testNeboBC.cpp:42-139
        key, pl, gl = ("Gradient", "SVol", "SSurfX"), "SVol", "SSurfX"

        def want(i, j):
            if (i, j) in ((-1, 1), (0, 2)):
                B, C = 25 + (i + 1) + j * 5, (i + 1) + j * 5
                return (B - C * (1 / DX)) / (-1 / DX)
            if (i, j) in ((2, 1), (3, 2)):
                B, A = 25 + i + j * 5, (i - 1) + j * 5
                return (B - A * (-1 / DX)) / (1 / DX)
            return i + j * 5
        minus, plus = [(0, 1, 0), (1, 2, 0)], [(2, 1, 0), (3, 2, 0)]
    elif name == "XVol GradX":        # :141-247
        key, pl, gl = ("Gradient", "XVol", "XSurfX"), "XVol", "XSurfX"

        def want(i, j):
            if (i, j) in ((-1, 1), (-1, 2)):
                B, C = 25 + i + j * 5, (i + 1) + j * 5
                return (B - C * (1 / DX)) / (-1 / DX)
            if (i, j) in ((4, 1), (4, 2)):
                B, A = 25 + (i - 1) + j * 5, (i - 1) + j * 5
                return (B - A * (-1 / DX)) / (1 / DX)
            return i + j * 5
        minus, plus = [(-1, 1, 0), (-1, 2, 0)], [(3, 1, 0), (3, 2, 0)]
    elif name == "YVol GradY":        # :249-354
        key, pl, gl = ("Gradient", "YVol", "YSurfY"), "YVol", "YSurfY"

        def want(i, j):
            if (i, j) in ((0, -1), (1, -1)):
                B, C = 25 + i + j * 5, i + (j + 1) * 5
                return (B - C * (1 / DY)) / (-1 / DY)
            if (i, j) in ((0, 4), (3, 4)):
                B, A = 25 + i + (j - 1) * 5, i + (j - 1) * 5
                return (B - A * (-1 / DY)) / (1 / DY)
            return i + j * 5
        minus, plus = [(0, -1, 0), (1, -1, 0)], [(0, 3, 0), (3, 3, 0)]
    elif name == "XVol FD GradX":     # :356-450
        key, pl, gl = ("GradientX", "XVol", "XVol"), "XVol", "XVol"

        def want(i, j):
            if (i, j) in ((-1, 0), (-1, 1)):
                B, C = 25 + (i + 1) + j * 5, (i + 2) + j * 5
                return (B - C * (0.5 / DX)) / (-0.5 / DX)
            if (i, j) in ((4, 0), (4, 1)):
                B, A = 25 + (i - 1) + j * 5, (i - 2) + j * 5
                return (B - A * (-0.5 / DX)) / (0.5 / DX)
            return i + j * 5
        minus, plus = [(0, 0, 0), (0, 1, 0)], [(3, 0, 0), (3, 1, 0)]
    else:
        raise KeyError(name)
    P = nh.Layout.build(DIM, 1, pl, BC)
    G = nh.Layout.build(DIM, 1, gl, BC)
    phi = _grid(P, lambda i, j: i + j * 5)
    gamma = _grid(G, lambda i, j: 25 + i + j * 5)
    expected = _grid(P, want)
    return db[key], P, G, phi, gamma, expected, minus, plus


CASES = ("SVol GradX", "XVol GradX", "YVol GradY", "XVol FD GradX")
