"""Expression trees of the Python binding and their serialisation to nebo_expr_t.

Nodes mirror the reference's expression-template nodes: scalars and field leaves
(spatialops/NeboRhs.h), pointwise operators (NeboOperators.h), stencils
(NeboStencils.h), cond clause chains (NeboCond.h) and masks (NeboMask.h).
``cond(t1, v1)(t2, v2)(default)`` picks the first true clause, as
NeboCond.h:408-412 does.
"""
from __future__ import annotations

import ctypes as C

from . import _abi
from .fields import Field, Mask

_BOOL_OPS = {"EQ", "NE", "LT", "LE", "GT", "GE", "AND", "OR", "NOT", "MASK", "MASK_SHIFT"}


class Expr:
    __slots__ = ("op", "kids", "field", "mask", "value", "stencil", "idx")
    __array_priority__ = 1000  # numpy scalars defer to Expr operators

    def __init__(self, op, kids=(), field=None, mask=None, value=0.0, stencil=None, idx=0):
        self.op = op
        self.kids = tuple(kids)
        self.field = field
        self.mask = mask
        self.value = value
        self.stencil = stencil
        self.idx = idx

    @property
    def boolean(self):
        return self.op in _BOOL_OPS

    # arithmetic
    def __add__(self, o): return Expr("ADD", (self, E(o)))
    def __radd__(self, o): return Expr("ADD", (E(o), self))
    def __sub__(self, o): return Expr("SUB", (self, E(o)))
    def __rsub__(self, o): return Expr("SUB", (E(o), self))
    def __mul__(self, o): return Expr("MUL", (self, E(o)))
    def __rmul__(self, o): return Expr("MUL", (E(o), self))
    def __truediv__(self, o): return Expr("DIV", (self, E(o)))
    def __rtruediv__(self, o): return Expr("DIV", (E(o), self))
    def __neg__(self): return Expr("NEG", (self,))
    # comparisons -> boolean expressions
    def __lt__(self, o): return Expr("LT", (self, E(o)))
    def __le__(self, o): return Expr("LE", (self, E(o)))
    def __gt__(self, o): return Expr("GT", (self, E(o)))
    def __ge__(self, o): return Expr("GE", (self, E(o)))
    def eq(self, o): return Expr("EQ", (self, E(o)))
    def ne(self, o): return Expr("NE", (self, E(o)))
    def __and__(self, o): return Expr("AND", (self, E(o)))
    def __or__(self, o): return Expr("OR", (self, E(o)))
    def __invert__(self): return Expr("NOT", (self,))
    __hash__ = object.__hash__


def E(x) -> Expr:
    """Lift a float, Field, Mask or Expr into an expression node."""
    if isinstance(x, Expr):
        return x
    if isinstance(x, Field):
        return Expr("FIELD", field=x)
    if isinstance(x, Mask):
        return Expr("MASK", mask=x)
    return Expr("SCALAR", value=float(x))


# pointwise functions (NeboOperators.h)
def _un(op):
    return lambda a: Expr(op, (E(a),))


sin, cos, tan, exp, tanh, log, log10 = (_un(o) for o in ("SIN", "COS", "TAN", "EXP", "TANH", "LOG", "LOG10"))
sqrt, square, cube, erf, erfc = (_un(o) for o in ("SQRT", "SQUARE", "CUBE", "ERF", "ERFC"))
inv_erf, inv_erfc = _un("INV_ERF"), _un("INV_ERFC")   # erf^-1, erfc^-1 (NeboOperators.h:6150-6330)
abs_ = _un("ABS")


def max_(a, b): return Expr("MAX", (E(a), E(b)))
def min_(a, b): return Expr("MIN", (E(a), E(b)))
def pow_(a, b): return Expr("POW", (E(a), E(b)))


def powint(a, n: int):
    """pow(expr, int) -> PowIntFcn (NeboOperators.h:14598-14710)."""
    return Expr("POWINT", (E(a),), idx=int(n))


class _Cond:
    """cond(test, value)(test, value)...(default) builder (NeboCond.h CondBuilder)."""

    def __init__(self, clauses):
        self.clauses = clauses

    def __call__(self, *args):
        if len(args) == 2:
            return _Cond(self.clauses + [(E(args[0]), E(args[1]))])
        (default,) = args
        node = E(default)
        for test, val in reversed(self.clauses):
            node = Expr("COND", (test, val, node))
        return node


def cond(test, value):
    return _Cond([(E(test), E(value))])


class Stencil:
    """A weighted (sum_p arg(x+p)*c_p), plain-sum or edgeless stencil over fixed
    points.  An edgeless stencil folds like a weighted one but does not shrink the
    valid range on the side it reads (NeboStencils.h:526-760)."""

    def __init__(self, points, coefs=None, kind=_abi.STENCIL_WEIGHTED, src="SVol", dst="SVol", name="",
                 axis=None):
        self.points = [tuple(int(v) for v in p) for p in points]
        self.coefs = [float(c) for c in coefs] if coefs is not None else [1.0] * len(self.points)
        self.kind = kind
        self.src = src
        self.dst = dst
        self.name = name
        self.axis = axis   # direction of a same-location (FD) operator
        assert 1 <= len(self.points) <= _abi.NEBO_MAX_STENCIL_POINTS
        assert len(self.coefs) == len(self.points)

    def __call__(self, arg) -> Expr:
        return Expr("STENCIL", (E(arg),), stencil=self)

    def struct(self) -> _abi.StencilT:
        s = _abi.StencilT()
        s.kind = self.kind
        s.npts = len(self.points)
        for i, (p, c) in enumerate(zip(self.points, self.coefs)):
            s.pts[i][:] = list(p)
            s.coefs[i] = c
        return s


class Serialized:
    """A nebo_expr_t and the ctypes arrays that keep it alive."""

    def __init__(self, expr_t, keep):
        self.expr = expr_t
        self._keep = keep

    @property
    def ref(self):
        return C.byref(self.expr)


def serialize(root: Expr, where: str = "gpu") -> Serialized:
    """Post-order node list (children before parents, root last) with shared
    field / mask / stencil tables.  ``where`` picks the host or GPU copies."""
    nodes, fields, masks, stencils = [], [], [], []
    fidx, midx, sidx, memo = {}, {}, {}, {}

    def visit(e: Expr) -> int:
        key = id(e)
        if key in memo:
            return memo[key]
        n = _abi.NodeT()
        n.op = _abi.OP[e.op]
        kids = [visit(k) for k in e.kids]
        n.a, n.b, n.c = (kids + [-1, -1, -1])[:3]
        if e.op == "FIELD":
            k = id(e.field)
            if k not in fidx:
                fidx[k] = len(fields)
                fields.append(e.field.desc(where))
            n.idx = fidx[k]
        elif e.op == "MASK":
            k = id(e.mask)
            if k not in midx:
                midx[k] = len(masks)
                masks.append(e.mask.desc(where))
            n.idx = midx[k]
        elif e.op in ("STENCIL", "MASK_SHIFT"):
            k = id(e.stencil)
            if k not in sidx:
                sidx[k] = len(stencils)
                stencils.append(e.stencil.struct())
            n.idx = sidx[k]
        elif e.op == "POWINT":
            n.idx = e.idx
        n.value = e.value
        nodes.append(n)
        memo[key] = len(nodes) - 1
        return memo[key]

    visit(E(root))
    arr_n = (_abi.NodeT * len(nodes))(*nodes)
    arr_f = (_abi.FieldT * max(1, len(fields)))(*fields)
    arr_s = (_abi.StencilT * max(1, len(stencils)))(*stencils)
    arr_m = (_abi.MaskT * max(1, len(masks)))(*masks)
    t = _abi.ExprT(len(nodes), arr_n, len(fields), arr_f, len(stencils), arr_s, len(masks), arr_m)
    return Serialized(t, (arr_n, arr_f, arr_s, arr_m, root))


def mask_shift(mask, point) -> Expr:
    """mask(x + point) as a boolean expression (NeboMaskShift,
    NeboStencils.h:1048-1213)."""
    m = E(mask)
    if m.op != "MASK":
        raise TypeError("mask_shift takes a Mask")
    return Expr("MASK_SHIFT", (m,), stencil=Stencil([tuple(point)], name="MaskShift"))


def _lift_methods(cls):
    """Let Field and Mask take part in expressions directly (phi + 1.0, ~mask)."""
    for name in ("__add__", "__radd__", "__sub__", "__rsub__", "__mul__", "__rmul__",
                 "__truediv__", "__rtruediv__", "__neg__", "__lt__", "__le__", "__gt__",
                 "__ge__", "eq", "ne", "__and__", "__or__", "__invert__"):
        def method(self, *args, _n=name):
            return getattr(E(self), _n)(*args)
        setattr(cls, name, method)
    cls.__hash__ = object.__hash__
    cls.__array_priority__ = 1000


_lift_methods(Field)
_lift_methods(Mask)
