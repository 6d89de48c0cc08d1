"""Boundary conditions on the GPU (NeboBoundaryConditionBuilder and
NeboMaskShiftBuilder, spatialops/NeboStencilBuilder.h:504-764), through the C-ABI.

Bar: bit-exact against the oracle (oracle_bc_apply, pinned in
tests/test_oracle_bc.py) and against the known answers of
spatialops/structured/test/testNeboBC.cpp.
"""
import random

import numpy as np
import pytest
import torch

import nebo_hip as nh
from nebo_hip import _abi
import oracle_lib as orc
from bc_cases import CASES, case
from test_oracle_bc import bc_operators, random_bc_problem

pytestmark = pytest.mark.gpu


def both(layout, data):
    return nh.Field.from_numpy(layout, data, where="both")


def fetch(f):
    torch.cuda.synchronize()
    return f.gpu.cpu().numpy()


@pytest.mark.parametrize("name", CASES)
def test_reference_cases_on_gpu(name):
    op, P, G, phi, gamma, expected, minus, plus = case(name)
    f, g = both(P, phi), both(G, gamma)
    bc = nh.BoundaryCondition(op)
    bc(nh.Mask(G, minus), f, g, True)
    bc(nh.Mask.build(f, G.loc, plus), f, g, False)
    assert np.array_equal(fetch(f), expected.reshape(-1))


_KEYS = bc_operators()[1]


@pytest.mark.parametrize("key", _KEYS)
def test_every_operator_matches_oracle(key):
    bcflags = (True, False, True) if _KEYS.index(key) % 2 else (False, True, False)
    op, P, G, phi, gamma, minus, plus = random_bc_problem(key, 5, bcflags, ghost=2 if "Vol" in key[1] else 1)
    bc = nh.BoundaryCondition(op)
    f, g = both(P, phi), both(G, gamma)
    for pts, side in ((minus, True), (plus, False)):
        orc.bc_apply(op, pts, f, g, side)
        bc(nh.Mask(G, pts), f, g, side)
    assert np.array_equal(fetch(f), f.host)


@pytest.mark.parametrize("key", [("Gradient", "SVol", "SSurfX"), ("Interpolant", "SVol", "YSurfZ"),
                                 ("GradientY", "ZVol", "ZVol"), ("Divergence", "XSurfX", "XVol")])
def test_gamma_expression_and_scalar(key):
    op, P, G, phi, gamma, minus, plus = random_bc_problem(key, 9)
    bc = nh.BoundaryCondition(op)
    f, g = both(P, phi), both(G, gamma)
    orc.bc_apply(op, minus, f, g * 2.0 + 1.0, True)
    bc(nh.Mask(G, minus), f, g * 2.0 + 1.0, True)
    orc.bc_apply(op, plus, f, -0.75, False)
    bc(nh.Mask(G, plus), f, -0.75, False)
    assert np.array_equal(fetch(f), f.host)


@pytest.mark.parametrize("key", [("Gradient", "SVol", "SSurfX"), ("Interpolant", "SVol", "XSurfY"),
                                 ("GradientZ", "YVol", "YVol"), ("Divergence", "SSurfY", "SVol"),
                                 ("Gradient", "XVol", "XSurfX")])
def test_cond_form_on_gpu(key):
    """out <<= cond(shift(mask), rhs)(phi) over the whole field: the reference's
    GPU form.  Its edgeless reads run past the blocks at unselected points; those
    reads are bounds-checked and the selected values equal the point form's."""
    op, P, G, phi, gamma, minus, plus = random_bc_problem(key, 4, (True, True, True))
    bc = nh.BoundaryCondition(op)
    for pts, side in ((minus, True), (plus, False)):
        m = nh.Mask(G, pts).add_consumer(0)
        src, g = both(P, phi), both(G, gamma)
        out = both(P, phi)
        nh.assign(out, bc.cond_expr(m, src, g, side))
        ref = both(P, phi)
        orc.assign(ref, bc.cond_expr(m, src, g, side))
        got = fetch(out)
        assert np.array_equal(got, ref.host)
        point = both(P, phi)
        orc.bc_apply(op, pts, point, g, side)
        assert np.array_equal(got, point.host)


@pytest.mark.parametrize("P", [(1, 0, 0), (0, -1, 0), (0, 0, 2), (-1, 1, -1)])
def test_mask_shift_matches_oracle(P):
    L = nh.Layout.build((9, 7, 6), 2, "SVol", (True, False, True))
    rng = random.Random(sum(P) + 10)
    pts = [(rng.randint(-2, 10), rng.randint(-2, 8), rng.randint(-2, 7)) for _ in range(60)]
    m = nh.Mask(L, pts).add_consumer(0)
    r = np.random.default_rng(3)
    f, g = both(L, r.uniform(-1, 1, L.npts)), both(L, r.uniform(-1, 1, L.npts))
    e = nh.cond(nh.mask_shift(m, P), f * 3.0)(nh.mask_shift(m, (0, 0, 0)) | nh.mask_shift(m, P), -1.0)(g)
    out, ref = both(L, np.zeros(L.npts)), both(L, np.zeros(L.npts))
    nh.assign(out, e)
    orc.assign(ref, e)
    assert np.array_equal(fetch(out), ref.host)
    assert nh.assign_limits(out, e) == orc.limits(ref, e)


def test_conflicting_points_are_refused_on_gpu():
    op, P, G, phi, gamma, expected, minus, plus = case("SVol GradX")
    f, g = both(P, phi), both(G, gamma)
    bc = nh.BoundaryCondition(op)
    # two minus points one cell apart in x: the first writes what the second reads
    with pytest.raises(nh.NeboError, match="order of the writes"):
        bc(nh.Mask(G, [(0, 1, 0), (1, 1, 0)]), f, g, True)
    assert np.array_equal(fetch(f), phi.reshape(-1))


def test_inplace_cond_form_is_refused():
    """phi <<= cond(shift(mask), rhs(phi))(phi) reads phi through stencils while
    writing it; the engine refuses it (the point form is the in-place path)."""
    op, P, G, phi, gamma, expected, minus, plus = case("SVol GradX")
    f, g = both(P, phi), both(G, gamma)
    bc = nh.BoundaryCondition(op)
    m = nh.Mask(G, minus).add_consumer(0)
    with pytest.raises(nh.NeboError, match="read by a stencil"):
        nh.assign(f, bc.cond_expr(m, f, g, True))
