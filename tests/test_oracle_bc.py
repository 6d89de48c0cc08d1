"""Boundary-condition inversion (NeboBoundaryConditionBuilder and
NeboMaskShiftBuilder, spatialops/NeboStencilBuilder.h:504-764) on the CPU.

Pinning of the oracle's restatement (oracle_bc_apply):
  * the four blocks of spatialops/structured/test/testNeboBC.cpp reproduce the
    test's own expected values bit for bit (tests/bc_cases.py);
  * on random 3-D data for every operator of build_stencils with two or four
    points, applying the operator to the result at each mask point gives gamma
    back to rounding: the inversion property that defines the condition.
Then the product's builder (nebo_hip.bc) is checked through the oracle's tree
walker: one pass over its shifted points of its edgeless expression equals the
oracle's direct restatement bit for bit.  The engine's ghost rules for mask
shifts and edgeless stencils are compared with the oracle and with closed
forms, and the engine's host-side checks of masked assignments (write/read
conflicts, edgeless reads outside a block) are exercised without a GPU.
"""
import ctypes as C
import random

import numpy as np
import pytest

import nebo_hip as nh
from nebo_hip import _abi
import oracle_lib as orc
from bc_cases import CASES, case


def host(layout, arr=None):
    f = nh.Field.empty(layout, "cpu")
    f.host[:] = 0.0 if arr is None else np.asarray(arr, dtype=np.float64).reshape(-1)
    return f


@pytest.mark.parametrize("name", CASES)
def test_reference_cases_oracle(name):
    op, P, G, phi, gamma, expected, minus, plus = case(name)
    f, g = host(P, phi), host(G, gamma)
    orc.bc_apply(op, minus, f, g, True)
    orc.bc_apply(op, plus, f, g, False)
    assert np.array_equal(f.host, expected.reshape(-1))


@pytest.mark.parametrize("name", CASES)
def test_reference_cases_builder_through_oracle_walker(name):
    op, P, G, phi, gamma, expected, minus, plus = case(name)
    f, g = host(P, phi), host(G, gamma)
    bc = nh.BoundaryCondition(op)
    m_minus = nh.Mask(G, minus)
    m_plus = nh.Mask.build(P, G.loc, plus)
    for m, side in ((m_minus, True), (m_plus, False)):
        orc.masked_assign_points(bc.points(m, side), f, bc.rhs(f, g, side), passes=1)
    assert np.array_equal(f.host, expected.reshape(-1))


def bc_operators():
    db = nh.build_stencils(6, 5, 4, 1.2, 1.0, 0.8)
    keys = []
    for key, op in db.items():
        if not isinstance(op, nh.Stencil) or op.kind != _abi.STENCIL_WEIGHTED:
            continue
        if len(op.points) < 2:
            continue
        keys.append(key)
    return db, sorted(keys)


def _in_block(layout, q):
    return all(0 <= layout.valid_minus[d] + q[d] < layout.glob[d] for d in range(3))


def pick_points(op, P, G, minus, rng, density=0.5):
    """Gamma points whose condition reads and writes stay inside the blocks, thinned
    so that no point reads a phi value that another point writes."""
    bc = nh.BoundaryCondition(op)
    shift = bc.shift.minus_point if minus else bc.shift.plus_point
    side = op.points[0] if minus else op.points[-1]
    lo = [-G.valid_minus[d] for d in range(3)]
    hi = [G.glob[d] - G.valid_minus[d] for d in range(3)]
    cands = [(x, y, z) for z in range(lo[2], hi[2]) for y in range(lo[1], hi[1]) for x in range(lo[0], hi[0])]
    rng.shuffle(cands)
    chosen, written, read = [], set(), set()
    for g in cands:
        if rng.random() > density:
            continue
        x = tuple(g[d] - shift[d] for d in range(3))
        reads = [tuple(x[d] + p[d] - side[d] for d in range(3)) for p in op.points if p != side]
        if not _in_block(P, x) or not all(_in_block(P, q) for q in reads):
            continue
        if not _in_block(G, tuple(x[d] - side[d] for d in range(3))):
            continue
        if x in written or x in read or any(q in written for q in reads):
            continue
        chosen.append(g)
        written.add(x)
        read.update(reads)
    return chosen


def random_bc_problem(key, seed, bcflags=(False, False, False), ghost=1):
    db, _ = bc_operators()
    op = db[key]
    rng = random.Random(seed)
    n = (6, 5, 4)
    P = nh.Layout.build(n, ghost, op.src, bcflags)
    G = nh.Layout.build(n, ghost, op.dst, bcflags)
    r = np.random.default_rng(seed)
    phi = r.uniform(-1, 1, P.npts)
    gamma = r.uniform(-1, 1, G.npts)
    minus = pick_points(op, P, G, True, rng)
    plus = pick_points(op, P, G, False, rng)
    return op, P, G, phi, gamma, minus, plus


def _keys():
    return bc_operators()[1]


@pytest.mark.parametrize("key", _keys())
def test_inversion_property(key):
    """op(phi)(g) == gamma(g) after the condition, to rounding."""
    op, P, G, phi, gamma, minus, plus = random_bc_problem(key, 7)
    for pts, side in ((minus, True), (plus, False)):
        if not pts:
            continue
        f, g = host(P, phi), host(G, gamma)
        orc.bc_apply(op, pts, f, g, side)
        s = nh.serialize(op(f), "host")
        scale = max(abs(c) for c in op.coefs)
        for q in pts:
            v = orc.lib().oracle_eval(s.ref, s.expr.n_nodes - 1, *q)
            gi = g.array()[G.valid_minus[2] + q[2], G.valid_minus[1] + q[1], G.valid_minus[0] + q[0]]
            assert abs(v - gi) <= 1e-13 * scale * 8, (q, v, gi)
        # and nothing but the shifted points changed
        bc = nh.BoundaryCondition(op)
        xs = bc.points(nh.Mask(G, pts), side)
        changed = np.flatnonzero(f.host != phi)
        flat = [(P.valid_minus[0] + x[0]) + P.glob[0] * ((P.valid_minus[1] + x[1]) + P.glob[1] * (P.valid_minus[2] + x[2]))
                for x in xs]
        assert set(changed.tolist()) <= set(flat)


@pytest.mark.parametrize("key", _keys())
@pytest.mark.parametrize("bcflags", [(False, False, False), (True, True, True)])
def test_builder_matches_oracle_restatement(key, bcflags):
    op, P, G, phi, gamma, minus, plus = random_bc_problem(key, 11, bcflags)
    bc = nh.BoundaryCondition(op)
    a, ga = host(P, phi), host(G, gamma)
    b, gb = host(P, phi), host(G, gamma)
    for pts, side in ((minus, True), (plus, False)):
        orc.bc_apply(op, pts, a, ga, side)
        if pts:
            orc.masked_assign_points(bc.points(nh.Mask(G, pts), side), b, bc.rhs(b, gb, side), passes=1)
    assert np.array_equal(a.host, b.host)


@pytest.mark.parametrize("key", [("Gradient", "SVol", "SSurfX"), ("Interpolant", "SVol", "XSurfY"),
                                 ("GradientZ", "YVol", "YVol"), ("Divergence", "SSurfY", "SVol")])
def test_cond_form_equals_point_form(key):
    """The reference's GPU form, lhs <<= cond(shift(mask), rhs)(phi) over every
    point, gives the point form's values (oracle walker for both)."""
    op, P, G, phi, gamma, minus, plus = random_bc_problem(key, 3)
    bc = nh.BoundaryCondition(op)
    for pts, side in ((minus, True), (plus, False)):
        a, b = host(P, phi), host(P, phi)
        g = host(G, gamma)
        orc.bc_apply(op, pts, a, g, side)
        m = nh.Mask(G, pts)
        out = host(P, phi)
        src = host(P, phi)
        orc.assign(out, bc.cond_expr(m, src, g, side))
        lim = orc.limits(out, bc.cond_expr(m, src, g, side))
        # every shifted point lies in the swept range, where the two forms agree
        for x in bc.points(m, side):
            assert all(lim[0][d] <= x[d] < lim[1][d] for d in range(3))
        assert np.array_equal(out.host, a.host)
        del b


# ---------------------------------------------------------------------------
# ghost rules
# ---------------------------------------------------------------------------
def _lib_ghost_limits(lhs, expr):
    s = nh.serialize(expr, "host")
    d = lhs.desc("host")
    lim = _abi.LimitsT()
    empty = C.c_int(0)
    rc = nh.lib().nebo_assign_limits(C.byref(d), s.ref, 1, C.byref(lim), C.byref(empty))
    assert rc == 0, nh.lib().nebo_last_error()
    return None if empty.value else (tuple(lim.lo), tuple(lim.hi))


@pytest.mark.parametrize("p", [(1, 0, 0), (-1, 0, 0), (0, 2, 0), (0, 0, -2), (1, -1, 1)])
def test_edgeless_and_mask_shift_ghosts_closed_form(p):
    n = (8, 8, 8)
    L = nh.Layout.build(n, 3)
    L1 = L.with_valid(1, 1)          # operands valid one layer out, the LHS three
    f, lhs = host(L1), host(L)
    ed = nh.Stencil([p], [1.0], kind=_abi.STENCIL_EDGELESS)
    # edgeless: the valid range widens by max(p, 0) on the minus side and by
    # max(-p, 0) on the plus side; it never shrinks (a weighted stencil would)
    e = ed(f)
    lo = tuple(-min(3, 1 + max(v, 0)) for v in p)
    hi = tuple(8 + min(3, 1 + max(-v, 0)) for v in p)
    assert _lib_ghost_limits(lhs, e) == (lo, hi) == orc.limits(lhs, e)
    w = nh.Stencil([p], [1.0])(f)
    lo_w = tuple(-min(3, 1 + v) for v in p)
    hi_w = tuple(8 + min(3, 1 - v) for v in p)
    assert _lib_ghost_limits(lhs, w) == (lo_w, hi_w) == orc.limits(lhs, w)
    # mask shift: ghosts + (p, -p), the weighted rule
    m = nh.Mask(L1, [(0, 0, 0)])
    e = nh.cond(nh.mask_shift(m, p), 1.0)(0.0)
    assert _lib_ghost_limits(lhs, e) == (lo_w, hi_w) == orc.limits(lhs, e)


@pytest.mark.parametrize("seed", range(12))
def test_ghosts_of_bc_trees_agree_with_oracle(seed):
    rng = random.Random(seed)
    _, keys = bc_operators()
    db, _ = bc_operators()
    key = rng.choice(keys)
    op = db[key]
    bcf = tuple(rng.random() < 0.5 for _ in range(3))
    g = rng.randint(1, 2)
    P = nh.Layout.build((5, 4, 3), g, op.src, bcf)
    G = nh.Layout.build((5, 4, 3), g, op.dst, bcf)
    bc = nh.BoundaryCondition(op)
    m = nh.Mask(G, [(0, 0, 0)])
    phi, gam, lhs = host(P), host(G), host(P)
    for side in (True, False):
        for e in (bc.rhs(phi, gam, side), bc.cond_expr(m, phi, gam, side)):
            assert _lib_ghost_limits(lhs, e) == orc.limits(lhs, e)


def test_mask_shift_needs_a_mask_and_one_point():
    L = nh.Layout.build((4, 4, 4), 1)
    f = host(L)
    with pytest.raises(TypeError):
        nh.mask_shift(f, (1, 0, 0))
    m = nh.Mask(L, [(0, 0, 0)])
    e = nh.mask_shift(m, (1, 0, 0))
    e.stencil.points.append((0, 1, 0))
    e.stencil.coefs.append(1.0)
    s = nh.serialize(nh.cond(e, 1.0)(0.0), "host")
    d = f.desc("host")
    rc = nh.lib().nebo_assign_limits(C.byref(d), s.ref, 1, C.byref(_abi.LimitsT()), C.byref(C.c_int()))
    assert rc == _abi.NEBO_ERR_INVALID
    assert "one-point" in nh.lib().nebo_last_error().decode()


# ---------------------------------------------------------------------------
# host checks of masked assignments (no GPU call is reached)
# ---------------------------------------------------------------------------
def _masked_rc(lhs, points, rhs):
    s = nh.serialize(rhs, "host")
    for i in range(s.expr.n_fields):
        s.expr.fields[i].device = 0
    d = lhs.layout.desc(lhs.host.ctypes.data, 0)
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.int32).reshape(-1))
    rc = nh.lib().nebo_masked_assign(C.byref(d), C.c_void_p(pts.ctypes.data), len(pts) // 3, s.ref, None)
    return rc, nh.lib().nebo_last_error().decode()


def test_masked_assign_refuses_write_read_conflicts():
    L = nh.Layout.build((6, 6, 6), 1)
    f = host(L)
    gx = nh.Stencil([(-1, 0, 0), (1, 0, 0)], [0.5, 0.5], kind=_abi.STENCIL_EDGELESS)
    # (1,1,1) reads (2,1,1), which the second point writes
    rc, msg = _masked_rc(f, [(1, 1, 1), (2, 1, 1)], gx(f))
    assert rc == _abi.NEBO_ERR_INVALID and "order of the writes" in msg
    # a repeated point while the LHS is read
    rc, msg = _masked_rc(f, [(1, 1, 1), (1, 1, 1)], f * 2.0)
    assert rc == _abi.NEBO_ERR_INVALID and "more than once" in msg


def test_masked_assign_refuses_edgeless_reads_outside_block():
    L = nh.Layout.build((6, 6, 6), 1)
    f, g = host(L), host(L)
    far = nh.Stencil([(3, 0, 0)], [1.0], kind=_abi.STENCIL_EDGELESS)
    rc, msg = _masked_rc(f, [(4, 1, 1)], far(g))
    assert rc == _abi.NEBO_ERR_GHOST and "outside the field's allocated block" in msg
