"""Boundary conditions: the mask shift and boundary-condition builders of
spatialops/NeboStencilBuilder.h:504-764.

A boundary condition inverts one stencil operator at a set of boundary points:
given an operator ``op`` from phi's location to gamma's location and a mask of
gamma points, it sets the one phi value the operator reads on the minus (or plus)
side of each mask point so that ``op(phi)`` equals ``gamma`` there.

Points of ``op`` are in definition order.  The low point is the first one (the
reference's ``PointCollection::Last`` with coefficient ``coefs().last()``), the
high point the last one (``PointCollection::First``, ``coefs().coef()``).  For a
mask point g, the phi point set is x = g - shift, where the shift of the minus
side is ``(gamma offset < phi offset)`` per component and that of the plus side is
``-(gamma offset > phi offset)`` (:504-516); a same-location (FD) operator shifts
by +e and -e along its direction (:518-526).  There
    phi(x) = (gamma(x - p_side) * 1.0 - fold_{i != side} phi(x + p_i - p_side) * c_i) / c_side
with both stencils edgeless (their valid range does not shrink).

The reference applies this in two ways.  On the CPU it loops once over the mask
points (``cpu_apply``, :646-680).  On a GPU it sweeps the whole field with
``phi <<= cond(shift(mask), rhs)(phi)`` (:712-719).  Both give the same values at
the mask points and leave every other point as it was.  :class:`BoundaryCondition`
runs the first form on the GPU: one masked assignment over the shifted points,
which touches only the boundary.  :meth:`BoundaryCondition.cond_expr` builds the
second form for callers that want it inside a larger expression.
"""
from __future__ import annotations

import numpy as np

from . import _abi
from .api import assign_points
from .expr import E, Expr, Stencil, cond, mask_shift
from .fields import LOCATIONS, Field, Mask


def shift_points(src: str, dst: str, axis=None):
    """(minus, plus) shift of NeboMaskShiftBuilder<Op, src, dst>
    (MaskShiftPoints, NeboStencilBuilder.h:504-526)."""
    if src == dst:
        if axis is None:
            raise ValueError("a mask shift between fields of one location needs the operator's direction")
        e = [0, 0, 0]
        e[axis] = 1
        return tuple(e), tuple(-v for v in e)
    so, do = LOCATIONS[src][0], LOCATIONS[dst][0]
    minus = tuple(1 if so[d] < do[d] else 0 for d in range(3))
    plus = tuple(-(1 if so[d] > do[d] else 0) for d in range(3))
    return minus, plus


class MaskShift:
    """NeboMaskShiftBuilder: turns a mask of ``src`` points into a boolean
    expression over ``dst`` points, mask(x + shift)."""

    def __init__(self, src: str, dst: str, axis=None):
        self.src, self.dst = src, dst
        self.minus_point, self.plus_point = shift_points(src, dst, axis)

    @staticmethod
    def for_operator(op: Stencil) -> "MaskShift":
        """The shift a boundary condition of ``op`` uses (gamma location to phi location)."""
        return MaskShift(op.dst, op.src, op.axis)

    def minus(self, mask: Mask) -> Expr:
        return mask_shift(mask, self.minus_point)

    def plus(self, mask: Mask) -> Expr:
        return mask_shift(mask, self.plus_point)


def _sub(p, q):
    return tuple(a - b for a, b in zip(p, q))


class BoundaryCondition:
    """NeboBoundaryConditionBuilder<Op>(op): ``bc(mask, phi, gamma, minus)``."""

    def __init__(self, op: Stencil):
        if op.kind == _abi.STENCIL_SUM or len(op.points) < 2:
            raise ValueError("a boundary condition inverts a weighted stencil of two or more points")
        self.op = op
        self.phi_loc, self.gamma_loc = op.src, op.dst
        self.shift = MaskShift.for_operator(op)
        pts, cs = op.points, op.coefs
        low, high = pts[0], pts[-1]
        self.low_coef, self.high_coef = cs[0], cs[-1]
        edg = _abi.STENCIL_EDGELESS
        g, p = self.gamma_loc, self.phi_loc
        self.minus_gamma = Stencil([_sub((0, 0, 0), low)], [1.0], edg, g, p, "MinusGamma")
        self.plus_gamma = Stencil([_sub((0, 0, 0), high)], [1.0], edg, g, p, "PlusGamma")
        self.minus_phi = Stencil([_sub(q, low) for q in pts[1:]], cs[1:], edg, p, p, "MinusPhi")
        self.plus_phi = Stencil([_sub(q, high) for q in pts[:-1]], cs[:-1], edg, p, p, "PlusPhi")

    def rhs(self, phi, gamma, minus: bool) -> Expr:
        """The inverted value at phi points: (gamma term - phi term) / coefficient."""
        if minus:
            return (self.minus_gamma(E(gamma)) - self.minus_phi(E(phi))) / self.low_coef
        return (self.plus_gamma(E(gamma)) - self.plus_phi(E(phi))) / self.high_coef

    def cond_expr(self, mask: Mask, phi, gamma, minus: bool) -> Expr:
        """cond(shift(mask), rhs)(phi): the reference's GPU form, over every point."""
        test = self.shift.minus(mask) if minus else self.shift.plus(mask)
        return cond(test, self.rhs(phi, gamma, minus))(phi)

    def points(self, mask: Mask, minus: bool) -> np.ndarray:
        """The phi points the condition sets: mask points minus the shift."""
        s = np.asarray(self.shift.minus_point if minus else self.shift.plus_point, dtype=np.int32)
        return np.asarray(mask.points, dtype=np.int32).reshape(-1, 3) - s

    def __call__(self, mask: Mask, phi: Field, gamma, minus: bool, stream=None) -> Field:
        return assign_points(self.points(mask, minus), phi, self.rhs(phi, gamma, minus), stream)
