"""Operator types and the operator database (the Python mirror of
spatialops/structured/stencil/FVStaggeredOperatorTypes.h, StencilBuilder.cpp and
OperatorDatabase.h).

Stencil points follow NeboStencilBuilder.h:154-232: for Stencil2 the low point is
-(src_offset > dest_offset) and the high point is (src_offset < dest_offset) per
component; Stencil4 combines a low/high pair in two directions; FD stencils use
-e_d and +e_d.  Coefficients follow build_stencils (StencilBuilder.cpp:60-72).
"""
from __future__ import annotations

import itertools

from . import _abi
from .expr import Expr, Stencil, E
from .fields import FACE_TYPES, LOCATIONS

VOLS = ("SVol", "XVol", "YVol", "ZVol")


def _off(loc):
    return LOCATIONS[loc][0]


def stencil2_points(src, dst):
    so, do = _off(src), _off(dst)
    low = tuple(-(1 if so[d] > do[d] else 0) for d in range(3))
    high = tuple(1 if so[d] < do[d] else 0 for d in range(3))
    return [low, high]


def stencil4_points(src, dst):
    so, do = _off(src), _off(dst)
    first = 0 if so[0] != do[0] else 1
    second = 2 if so[2] != do[2] else 1

    def lohi(d):
        lo = -(1 if so[d] > do[d] else 0)
        hi = 1 if so[d] < do[d] else 0
        return lo, hi

    lf, hf = lohi(first)
    ls, hs = lohi(second)

    def pt(a, b):
        p = [0, 0, 0]
        p[first] += a
        p[second] += b
        return tuple(p)

    return [pt(lf, ls), pt(hf, ls), pt(lf, hs), pt(hf, hs)]


def fd_points(axis):
    lo, hi = [0, 0, 0], [0, 0, 0]
    lo[axis], hi[axis] = -1, 1
    return [tuple(lo), tuple(hi)]


BOX27 = [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]


class Average:
    """NeboAverageStencilBuilder: sum stencil times 1/npts (NeboStencilBuilder.h:389-452)."""

    def __init__(self, points, src, dst):
        self.sum = Stencil(points, kind=_abi.STENCIL_SUM, src=src, dst=dst, name="BoxFilter")
        self.scale = 1.0 / float(len(points))
        self.src, self.dst = src, dst

    def __call__(self, arg) -> Expr:
        return Expr("MUL", (self.sum(arg), E(self.scale)))


class OperatorDatabase(dict):
    """(op type, source location, destination location) -> operator."""

    def retrieve_operator(self, op, src, dst):
        return self[(op, src, dst)]


def build_stencils(nx, ny, nz, Lx, Ly, Lz) -> OperatorDatabase:
    """build_stencils(grid, opDB) of StencilBuilder.cpp:52-220."""
    dx, dy, dz = Lx / nx, Ly / ny, Lz / nz
    half = (0.5, 0.5)
    dcoef = {0: (-1.0 / dx, 1.0 / dx), 1: (-1.0 / dy, 1.0 / dy), 2: (-1.0 / dz, 1.0 / dz)}
    hcoef = {0: (-0.5 / dx, 0.5 / dx), 1: (-0.5 / dy, 0.5 / dy), 2: (-0.5 / dz, 0.5 / dz)}
    db = OperatorDatabase()

    def s2(op, src, dst, coefs):
        db[(op, src, dst)] = Stencil(stencil2_points(src, dst), coefs, src=src, dst=dst, name=op)

    for vol in VOLS:
        for d, face in enumerate(FACE_TYPES[vol]):
            s2("Interpolant", vol, face, half)
            s2("Interpolant", face, vol, half)
            s2("Gradient", vol, face, dcoef[d])
            s2("Divergence", face, vol, dcoef[d])
    for (src, dst, d) in (("XVol", "YSurfX", 1), ("XVol", "ZSurfX", 2), ("YVol", "XSurfY", 0),
                          ("YVol", "ZSurfY", 2), ("ZVol", "XSurfZ", 0), ("ZVol", "YSurfZ", 1)):
        s2("Interpolant", src, dst, half)
        s2("Gradient", src, dst, dcoef[d])
    for vol, d in (("XVol", 0), ("YVol", 1), ("ZVol", 2)):
        s2("Interpolant", "SVol", vol, half)
        s2("Gradient", "SVol", vol, dcoef[d])
        s2("Interpolant", vol, "SVol", half)
        s2("Gradient", vol, "SVol", dcoef[d])
    null = [(0, 0, 0)]
    for src, dst in (("SVol", "SVol"), ("XVol", "XVol"), ("YVol", "YVol"), ("ZVol", "ZVol"),
                     ("XVol", "SSurfX"), ("YVol", "SSurfY"), ("ZVol", "SSurfZ"),
                     ("SVol", "XSurfX"), ("SVol", "YSurfY"), ("SVol", "ZSurfZ"),
                     ("XSurfX", "SVol"), ("YSurfY", "SVol"), ("ZSurfZ", "SVol")):
        db[("Interpolant", src, dst)] = Stencil(null, kind=_abi.STENCIL_SUM, src=src, dst=dst, name="Null")
    for src, dst in (("SVol", "XSurfY"), ("SVol", "XSurfZ"), ("SVol", "YSurfX"), ("SVol", "YSurfZ"),
                     ("SVol", "ZSurfX"), ("SVol", "ZSurfY"), ("XSurfY", "SVol"), ("XSurfZ", "SVol"),
                     ("YSurfX", "SVol"), ("YSurfZ", "SVol"), ("ZSurfX", "SVol"), ("ZSurfY", "SVol"),
                     ("XVol", "YVol"), ("XVol", "ZVol"), ("YVol", "XVol"), ("YVol", "ZVol"),
                     ("ZVol", "XVol"), ("ZVol", "YVol")):
        db[("Interpolant", src, dst)] = Stencil(stencil4_points(src, dst), (0.25,) * 4, src=src, dst=dst,
                                                name="Stencil4")
    for loc in VOLS + tuple(itertools.chain.from_iterable(FACE_TYPES[v] for v in VOLS[1:])):
        db[("Filter", loc, loc)] = Average(BOX27, loc, loc)
    for vol in VOLS:
        for d, ax in enumerate("XYZ"):
            db[(f"Interpolant{ax}", vol, vol)] = Stencil(fd_points(d), half, src=vol, dst=vol, name="FD",
                                                         axis=d)
            db[(f"Gradient{ax}", vol, vol)] = Stencil(fd_points(d), hcoef[d], src=vol, dst=vol, name="FD",
                                                      axis=d)
    return db


def basic_op_types(db: OperatorDatabase, vol="SVol") -> dict:
    """BasicOpTypes<Vol>: InterpC2F{X,Y,Z}, InterpF2C{X,Y,Z}, Grad{X,Y,Z}, Div{X,Y,Z}."""
    fx, fy, fz = FACE_TYPES[vol]
    out = {}
    for ax, face in zip("XYZ", (fx, fy, fz)):
        out[f"InterpC2F{ax}"] = db[("Interpolant", vol, face)]
        out[f"InterpF2C{ax}"] = db[("Interpolant", face, vol)]
        out[f"Grad{ax}"] = db[("Gradient", vol, face)]
        out[f"Div{ax}"] = db[("Divergence", face, vol)]
    return out


def axis_coefs(db: OperatorDatabase, vol="SVol") -> _abi.AxisCoefsT:
    """The per-axis coefficients of the basic operators as a nebo_axis_coefs_t."""
    ops = basic_op_types(db, vol)
    c = _abi.AxisCoefsT()
    for d, ax in enumerate("XYZ"):
        c.interp[d][:] = ops[f"InterpC2F{ax}"].coefs
        c.grad[d][:] = ops[f"Grad{ax}"].coefs
        c.div[d][:] = ops[f"Div{ax}"].coefs
    return c
