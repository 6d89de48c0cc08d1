"""GPU parity: the HIP engine against the CPU oracle on identical inputs.

Bar: bit-exact (np.array_equal on float64) for every stencil / pointwise / cond /
mask path; reductions within the stated bounds; transcendental ops within the ulp
budgets of the reference's test/NeboTest.cpp:497-523.  Every check runs through the
C-ABI (include/nebo_hip.h) via the nebo_hip ctypes binding.
"""
import math

import numpy as np
import pytest
import torch

import nebo_hip as nh
from nebo_hip import inputs, _abi
import oracle_lib as orc

pytestmark = pytest.mark.gpu


def both(layout, data):
    return nh.Field.from_numpy(layout, data, where="both")


def gpu_empty_like(layout, fill=0.0):
    f = nh.Field.empty(layout, "gpu")
    f.gpu.fill_(fill)
    f.host = np.full(layout.npts, fill)
    return f


def fetch(f):
    torch.cuda.synchronize()
    return f.gpu.cpu().numpy()


def _ops(n, L=(1.0, 1.0, 1.0)):
    db = nh.build_stencils(*n, *L)
    return db, nh.basic_op_types(db)


def _divkgrad_tree(ops, k, phi, form):
    D = [ops[f"Div{a}"](ops[f"InterpC2F{a}"](k) * ops[f"Grad{a}"](phi)) for a in "XYZ"]
    if form == 0:
        return D[0] + D[1] + D[2]
    return -D[0] - D[1] - D[2]


def test_device_visible():
    assert nh.device_count() >= 1
    nh.lib()
    assert torch.cuda.is_available()


# --------------------------------------------------------------------------
# var-coefficient diffusion (hand-written kernel and generic evaluator)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [(1, 1, 1), (2, 3, 5), (17, 9, 13), (64, 8, 33), (70, 66, 20)])
@pytest.mark.parametrize("form", [0, 1])
def test_divkgrad_small_bit_exact(n, form):
    print(f"divkgrad n={n} form={form}", flush=True)
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(sum(n) * 7 + form)
    k = both(L, rng.uniform(0.5, 2.0, L.npts))
    phi = both(L, rng.standard_normal(L.npts))
    db, ops = _ops(n)
    tree = _divkgrad_tree(ops, k, phi, form)
    ref = nh.Field.from_numpy(L, np.full(L.npts, 3.0), where="cpu")
    orc.assign(ref, tree)
    for generic in (False, True):
        r = gpu_empty_like(L, 3.0)
        nh.assign(r, tree, generic=generic)
        assert np.array_equal(fetch(r), ref.host), f"generic={generic}"
    r = gpu_empty_like(L, 3.0)
    nh.divkgrad(r, k, phi, nh.axis_coefs(db), form)
    assert np.array_equal(fetch(r), ref.host)


def test_divkgrad_256_bit_exact():
    """The headline workload: 256^3 fp64, seeded k ~ U[0.5,2), 8-mode Fourier phi,
    periodic ghosts.  The GPU result equals the oracle's closed form bit for bit."""
    n = 256
    print("divkgrad 256^3", flush=True)
    L = nh.Layout.build((n, n, n), 1)
    k = both(L, inputs.coefficient_field(n, 2))
    phi = both(L, inputs.fourier_field(n, 3))
    db, ops = _ops((n, n, n))
    c = nh.axis_coefs(db)
    ref = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.divkgrad(ref, k, phi, c, 0, threads=8)
    r = gpu_empty_like(L, 0.0)
    nh.assign(r, _divkgrad_tree(ops, k, phi, 0))
    out = fetch(r)
    assert np.array_equal(out, ref.host)
    # size-independent property: periodic divergence telescopes to ~0
    inner = out.reshape(L.glob[::-1])[1:-1, 1:-1, 1:-1]
    assert abs(math.fsum(inner.ravel())) < 1e-9 * np.abs(inner).sum()


def test_lapl_euler_bit_exact():
    n = (33, 17, 9)
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(4)
    phi = both(L, rng.standard_normal(L.npts))
    db, ops = _ops(n)
    s = 0.1 / (33.0 ** 2)
    lap = ops["DivX"](ops["GradX"](phi)) + ops["DivY"](ops["GradY"](phi)) + ops["DivZ"](ops["GradZ"](phi))
    ref = nh.Field.from_numpy(L, np.full(L.npts, -1.0), where="cpu")
    orc.assign(ref, phi + s * lap)
    assert nh.assign_plan(gpu_empty_like(L), phi + s * lap) == _abi.KERNEL_LAPL_EULER
    for generic in (False, True):
        out = gpu_empty_like(L, -1.0)
        nh.assign(out, phi + s * lap, generic=generic)
        assert np.array_equal(fetch(out), ref.host)


# --------------------------------------------------------------------------
# generic evaluator: pointwise, cond, masks, other stencil families
# --------------------------------------------------------------------------

def test_cond_pointwise_bit_exact():
    """c <<= cond(phi > 0, phi*phi)(-phi) at 64^3 (the pointwise part of the workload)."""
    n = 64
    L = nh.Layout.build((n, n, n), 1)
    phi = both(L, inputs.fourier_field(n, 3))
    e = nh.cond(phi > 0.0, phi * phi)(-phi)
    ref = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.assign(ref, e)
    c = gpu_empty_like(L)
    nh.assign(c, e)
    assert np.array_equal(fetch(c), ref.host)


def test_mask_known_answer_on_gpu():
    """spatialops/structured/test/testMask.cpp on the GPU: cond and masked_assign."""
    L = nh.Layout.build((28, 28, 28), 1)
    center = [(3, 3, 0), (4, 3, 0), (3, 4, 0), (4, 4, 0)]
    corners = [(1, 1, 0), (6, 1, 0), (1, 6, 0), (6, 6, 0)]
    expected = np.full(L.glob[::-1], 4.0)
    for (x, y, z) in center:
        expected[z + 1, y + 1, x + 1] = 3.0
    for (x, y, z) in corners:
        expected[z + 1, y + 1, x + 1] = 7.0
    m1 = nh.Mask(L, center).add_consumer(0)
    m2 = nh.Mask(L, corners).add_consumer(0)
    f = gpu_empty_like(L, -1.0)
    nh.assign(f, nh.cond(m1, 3.0)(m2, 7.0)(4.0))
    assert np.array_equal(fetch(f).reshape(L.glob[::-1]), expected)
    g = gpu_empty_like(L, -1.0)
    nh.assign(g, nh.E(4.0))
    nh.masked_assign(m1, g, 3.0)
    nh.masked_assign(m2, g, 7.0)
    assert np.array_equal(fetch(g).reshape(L.glob[::-1]), expected)


def test_heat_equation_mask_step():
    """examples/heat_equation.cpp pattern on a 2-D (nz = 1) domain with a +BC.
    Its fields are one plane each and may be allocated back to back, so a read
    of one field can lie one plane from a read of another: the column schedule
    must still keep each field's reads inside that field's allocation."""
    n = (6, 6, 1)
    L = nh.Layout.build(n, 1, "SVol", (True, True, True))
    left = [(-1, i, 0) for i in range(-1, n[1] + 1)]
    right = [(n[0], i, 0) for i in range(-1, n[1] + 1)]
    top = [(i, -1, 0) for i in range(-1, n[0] + 1)]
    bottom = [(i, n[1], 0) for i in range(-1, n[0] + 1)]
    ml, mr, mt, mb = (nh.Mask(L, p).add_consumer(0) for p in (left, right, top, bottom))
    phi = gpu_empty_like(L, 0.0)
    href = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    alpha = gpu_empty_like(L, 1.0)
    aref = nh.Field.from_numpy(L, np.ones(L.npts), where="cpu")
    db, _ = _ops((6, 6, 1))
    ops = nh.basic_op_types(db)
    rhs, rref = gpu_empty_like(L), nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    e0 = nh.cond(ml, 10.0)(mr, 0.0)(5.0)
    nh.assign(phi, e0)
    orc.assign(href, nh.cond(ml, 10.0)(mr, 0.0)(5.0))
    dt = 0.25 * (1 / 36) * (1 / 36) / ((1 / 36 + 1 / 36) * 1.0)
    for _ in range(5):
        for (P, A, R, where) in ((phi, alpha, rhs, "gpu"), (href, aref, rref, "cpu")):
            e = (ops["DivX"](ops["InterpC2FX"](A) * ops["GradX"](P))
                 + ops["DivY"](ops["InterpC2FY"](A) * ops["GradY"](P)))
            e2 = P + dt * R
            e3 = nh.cond(ml, 10.0)(mr, 0.0)(mt | mb, 5.0)(P)
            if where == "gpu":
                nh.assign(R, e)
                nh.assign(P, e2)
                nh.assign(P, e3)
            else:
                orc.assign(R, e)
                orc.assign(P, e2)
                orc.assign(P, e3)
        assert np.array_equal(fetch(phi), href.host)


@pytest.fixture(params=[True, False], ids=["compiled", "interpreter"])
def fused_mode(request):
    """Run a test on the kernels compiled at run time and on the interpreter."""
    nh.set_fused_jit(request.param)
    yield request.param
    nh.set_fused_jit(True)


@pytest.mark.parametrize("seed", range(12))
def test_random_trees_bit_exact(seed, fused_mode):
    rng = np.random.default_rng(seed)
    n = tuple(int(v) for v in rng.integers(3, 20, 3))
    L = nh.Layout.build(n, 2)
    db, ops = _ops(n, (1.0, 2.0, 0.5))
    a = both(L, rng.standard_normal(L.npts))
    b = both(L, rng.uniform(0.5, 2.0, L.npts))
    m = nh.Mask(L, rng.integers(-1, 3, (20, 3))).add_consumer(0)
    s4 = db[("Interpolant", "SVol", "XSurfY")]
    box = db[("Filter", "SVol", "SVol")]
    fdx = db[("GradientX", "SVol", "SVol")]
    exprs = [
        nh.max_(a, b) * nh.min_(a, 1.5) - nh.abs_(a) / b,
        nh.cond(a > b, a)(m, b * 2.0)(nh.sqrt(b)),
        s4(a * b) + box(a) - fdx(fdx(b)),
        ops["DivZ"](ops["GradZ"](a)) * nh.square(b) + nh.cube(a),
        nh.cond((a < 0.0) & ~(b > 1.0), -a)((a > 1.0) | m, 2.0)(a + b),
        nh.powint(a, 2) + nh.powint(b, 1),
    ]
    for i, e in enumerate(exprs):
        print(f"tree seed={seed} expr={i}", flush=True)
        ref = nh.Field.from_numpy(L, np.full(L.npts, 9.0), where="cpu")
        orc.assign(ref, e)
        out = gpu_empty_like(L, 9.0)
        nh.assign(out, e)
        got = fetch(out)
        if i == 5:
            # pow(x, 2): the host pow is within one ulp of x*x; the following add
            # can round that difference up to two ulps of the sum
            ulp = np.abs(got - ref.host) / np.spacing(np.maximum(np.abs(ref.host), 1e-300))
            assert ulp.max() <= 2.0
        else:
            assert np.array_equal(got, ref.host)


@pytest.fixture(params=[((0, -1, 0, 0, -1), (1, 0, 0)), ((0, -1, 0, 0, -1), (0, 0, 0)),
                        ((1024, 0, 8, 64, 1), (0, 0, 0)), ((64, -1, 16, 256, 0), (0, 0, 0)),
                        ((512, 64, 4, 48, 1), (0, 0, 0)), ((0, -1, 0, 0, -1), (1, 2, 7)),
                        ((0, -1, 0, 0, 0), (1, 4, 1)), ((0, -1, 0, 0, -1), (2, 1, 3))],
                ids=["default", "flat", "b1024-v8-xcd", "b64-v16", "tile64x8-xcd", "column-w2-z7",
                     "column-w4-z1", "one-column-z3"])
def jit_shape(request):
    """Compiled kernels of several launch shapes: the column schedule (rows and
    planes past the box end inside a block, 1-4 waves, 1-8 planes per block) and,
    with it off, flattened-plane and 2-D tiles, with and without the XCD
    renumbering, cells per thread from 1 to 16 (full chunks with shared z-chain
    reads, and tails cut by the box end)."""
    tune, column = request.param
    nh.fused_jit_tune(*tune)
    nh.fused_jit_column(*column)
    yield request.param
    nh.fused_jit_tune()
    nh.fused_jit_column()


@pytest.mark.parametrize("column", [(1, 0, 0), (1, 1, 3), (1, 3, 5), (1, 2, 64), (2, 0, 0), (2, 3, 5)],
                         ids=["default", "w1-z3", "w3-z5", "w2-z64", "one-column", "one-column-w3-z5"])
@pytest.mark.parametrize("n", [(100, 37, 21), (131, 9, 10)])
def test_column_schedule_matches_oracle(column, n):
    """Generic stencil trees on boxes with 2 ghost layers run on the column
    schedule, two columns per lane where the ring allows (mode 1) or one (mode
    2): wave columns in x partly past the box (100 or 131 wide: a lane whose
    second column lies past the box stores its first alone; ghost-inclusive
    widths 102 and 133), row groups and z chunks cut by the box end, 1-3 waves
    per block.  Results equal the oracle's bit for bit, with the interior
    assigned (ghost cells untouched) and with the ghost-inclusive region."""
    L = nh.Layout.build(n, 2)
    rng = np.random.default_rng(11)
    k = both(L, rng.uniform(0.5, 2.0, L.npts))
    phi = both(L, rng.standard_normal(L.npts))
    _, ops = _ops(n)
    div = sum((ops[f"Div{a}"](ops[f"InterpC2F{a}"](k) * ops[f"Grad{a}"](phi)) for a in "YZ"),
              ops["DivX"](ops["InterpC2FX"](k) * ops["GradX"](phi)))
    lap = sum(ops[f"Div{a}"](ops[f"Grad{a}"](phi)) for a in "YZ") + ops["DivX"](ops["GradX"](phi))
    mixed = k * lap + 0.5 * phi
    try:
        nh.fused_jit_column(*column)
        for name, t in (("div(k grad phi)", div), ("laplacian", lap), ("k*lap + phi/2", mixed)):
            for use_ghost in (True, False):
                print(f"column {column} {name} use_ghost={use_ghost}", flush=True)
                out = gpu_empty_like(L, 7.0)
                ref = nh.Field.from_numpy(L, np.full(L.npts, 7.0), where="cpu")
                nh.assign(out, t, use_ghost=use_ghost, generic=True)
                orc.assign(ref, t, use_ghost=use_ghost)
                assert np.array_equal(fetch(out), ref.host), (column, name, use_ghost)
    finally:
        nh.fused_jit_column()


def test_compiled_kernels_equal_interpreter(jit_shape):
    """Random trees over every opcode (transcendentals, pow, masks, mask shifts,
    edgeless stencils): the compiled kernel and the interpreter agree bit for bit,
    and every tree compiled."""
    import random
    from test_fused_source import random_tree
    n = (13, 9, 11)
    L = nh.Layout.build(n, 2)
    rng = np.random.default_rng(5)
    f = both(L, rng.uniform(0.2, 2.0, L.npts))
    g = both(L, rng.standard_normal(L.npts))
    db = nh.build_stencils(*n, 1.0, 1.0, 1.0)
    m = nh.Mask(L, rng.integers(-1, 5, (40, 3))).add_consumer(0)
    stencils = [db[("Interpolant", "SVol", "XSurfY")], db[("GradientY", "SVol", "SVol")],
                nh.Stencil([(1, 0, 0)], [2.0], kind=_abi.STENCIL_EDGELESS)]
    r = random.Random(3)
    c0, h0, f0 = nh.fused_jit_stats()
    trees = [random_tree(r, [f, g], stencils, [m, nh.mask_shift(m, (0, 1, 0))], 4) for _ in range(24)]
    ran = 0
    for i, t in enumerate(trees):
        print(f"compiled-vs-interpreter tree {i}", flush=True)
        try:
            nh.lower_stats(t, "gpu")
        except nh.NeboError:
            continue
        outs = []
        for jit in (True, False):
            nh.set_fused_jit(jit)
            out = gpu_empty_like(L, 7.0)
            try:
                nh.assign(out, t)
            except nh.NeboError:
                out = None
            outs.append(None if out is None else fetch(out))
        nh.set_fused_jit(True)
        assert (outs[0] is None) == (outs[1] is None)
        if outs[0] is not None:
            assert np.array_equal(outs[0], outs[1], equal_nan=True), i
            ran += 1
    c1, h1, f1 = nh.fused_jit_stats()
    # every tree ran on a compiled kernel: compiled now or found in the cache
    assert ran >= 12 and f1 == f0 and (c1 + h1) - (c0 + h0) >= ran


def test_compiled_reductions_match_interpreter(jit_shape):
    """Reductions of random trees on the compiled kernels and on the interpreter:
    max and min equal; sum and norm within one rounding of each other's
    double-double result (both fold cells into (hi, lo) pairs, in different
    orders); every reduction kernel compiled."""
    import random
    from test_fused_source import random_tree
    n = (17, 9, 11)
    L = nh.Layout.build(n, 2)
    rng = np.random.default_rng(6)
    f = both(L, rng.uniform(0.2, 2.0, L.npts))
    g = both(L, rng.standard_normal(L.npts))
    db = nh.build_stencils(*n, 1.0, 1.0, 1.0)
    m = nh.Mask(L, rng.integers(-1, 5, (40, 3))).add_consumer(0)
    stencils = [db[("Interpolant", "SVol", "XSurfY")], db[("GradientY", "SVol", "SVol")],
                nh.Stencil([(1, 0, 0)], [2.0], kind=_abi.STENCIL_EDGELESS)]
    r = random.Random(8)
    c0, h0, f0 = nh.fused_jit_stats()
    trees = [nh.E(g), f * g + 1.0, stencils[1](g) * f]
    trees += [random_tree(r, [f, g], stencils, [m, nh.mask_shift(m, (0, 1, 0))], 3) for _ in range(14)]
    ran = empty = 0
    for i, t in enumerate(trees):
        print(f"compiled-vs-interpreter reductions, tree {i}", flush=True)
        try:
            nh.lower_stats(t, "gpu")
        except nh.NeboError:
            continue
        for use_ghost in (True, False):
            got = {}
            for jit in (True, False):
                nh.set_fused_jit(jit)
                try:
                    got[jit] = [fn(t) for fn in ((nh.field_max, nh.field_min, nh.field_sum, nh.field_norm)
                                                 if use_ghost else
                                                 (nh.field_max_interior, nh.field_min_interior,
                                                  nh.field_sum_interior, nh.field_norm_interior))]
                except nh.NeboError:
                    got[jit] = None
            nh.set_fused_jit(True)
            assert (got[True] is None) == (got[False] is None)
            if got[True] is None:
                continue
            (mx, mn, sm, nr), (mx0, mn0, sm0, nr0) = got[True], got[False]
            same = lambda a, b: a == b or (math.isnan(a) and math.isnan(b))
            assert same(mx, mx0) and same(mn, mn0), i
            if math.isfinite(sm0):
                assert abs(sm - sm0) <= 2.0 ** -51 * abs(sm0) + 2.0 ** -90 * (abs(mx0) + abs(mn0)) * L.npts, i
            else:
                assert same(sm, sm0), i
            if math.isfinite(nr0):
                assert abs(nr - nr0) <= 4 * 2.0 ** -53 * nr0, i
            else:
                assert same(nr, nr0), i
            ran += 1
            # an empty region (a tree without fields, interior) launches no kernel
            empty += (mx0, mn0, sm0, nr0) == (-math.inf, math.inf, 0.0, 0.0)
    c1, h1, f1 = nh.fused_jit_stats()
    # every other reduction ran on a compiled kernel: compiled now or found in the cache
    assert ran >= 12 and f1 == f0 and (c1 + h1) - (c0 + h0) >= 4 * (ran - empty)


def _all_operators():
    db = nh.build_stencils(9, 7, 6, 1.0, 0.9, 1.3)
    return db, sorted(db.keys())


@pytest.mark.parametrize("bc", [(False, False, False), (True, True, True)])
def test_every_operator_bit_exact(bc):
    """Every operator build_stencils registers (two- and four-point staggered
    operators, FD, null interpolants, box filters), applied to a field of its
    source location and assigned to a field of its destination location, equals
    the oracle bit for bit (StencilBuilder.cpp:52-220 is the list)."""
    db, keys = _all_operators()
    rng = np.random.default_rng(12)
    for key in keys:
        op = db[key]
        src = nh.Layout.build((9, 7, 6), 1, op.src, bc)
        dst = nh.Layout.build((9, 7, 6), 1, op.dst, bc)
        f = both(src, rng.standard_normal(src.npts))
        r = gpu_empty_like(dst, 0.0)
        ref = nh.Field.from_numpy(dst, np.zeros(dst.npts), where="cpu")
        try:
            lim = orc.limits(ref, op(f))
        except RuntimeError:
            lim = "error"
        if lim == "error":
            with pytest.raises(nh.NeboError):
                nh.assign(r, op(f))
            continue
        nh.assign(r, op(f))
        orc.assign(ref, op(f))
        assert np.array_equal(fetch(r), ref.host), key


@pytest.mark.parametrize("lengths", [(1.0, 1.0, 1.0), (0.7, 0.37, 0.45)])
def test_every_stencil_tile_bit_exact(lengths):
    """The tile shape changes only the schedule: every instantiated shape gives the
    same bits, on a box that is not a multiple of any tile, with one spacing per
    axis and with one spacing for all axes."""
    n = (70, 37, 45)
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(8)
    k = both(L, rng.uniform(0.5, 2.0, L.npts))
    phi = both(L, rng.standard_normal(L.npts))
    db, ops = _ops(n, lengths)
    c = nh.axis_coefs(db)
    ref = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.divkgrad(ref, k, phi, c, 0)
    ref1 = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.divkgrad(ref1, k, phi, c, 1)
    lref = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.lapl_euler(lref, phi, c, 0.01)
    try:
        for tile in nh.stencil_tiles():
            nh.stencil_set_tile(0, *tile)
            nh.stencil_set_tile(1, *tile)
            for form in (0, 1):
                r = gpu_empty_like(L, 0.0)
                nh.divkgrad(r, k, phi, c, form)
                if form == 0:
                    assert np.array_equal(fetch(r), ref.host), tile
                else:
                    assert np.array_equal(fetch(r), ref1.host), tile
            o = gpu_empty_like(L, 0.0)
            nh.lapl_euler(o, phi, c, 0.01)
            assert np.array_equal(fetch(o), lref.host), tile
        best, times = nh.stencil_autotune(gpu_empty_like(L), k, phi, c)
        assert best in nh.stencil_tiles() and len(times) == len(nh.stencil_tiles())
        best, times = nh.stencil_autotune(gpu_empty_like(L), None, phi, c, kernel=nh.KERNEL_LAPL)
        assert best in nh.stencil_tiles() and nh.stencil_get_tile(1) == best
    finally:
        nh.stencil_reset_tile(0)
        nh.stencil_reset_tile(1)


def _two_ghost_case(n, seed):
    L = nh.Layout.build(n, 2)
    rng = np.random.default_rng(seed)
    k = both(L, rng.uniform(0.5, 2.0, L.npts))
    phi = both(L, rng.standard_normal(L.npts))
    _, ops = _ops(n)
    s = 0.1 / (n[0] ** 2)
    lap = ops["DivX"](ops["GradX"](phi)) + ops["DivY"](ops["GradY"](phi)) + ops["DivZ"](ops["GradZ"](phi))
    trees = {"divk form 0": _divkgrad_tree(ops, k, phi, 0), "divk form 1": _divkgrad_tree(ops, k, phi, 1),
             "lapl+euler": phi + s * lap}
    refs = {}
    for name, t in trees.items():
        for use_ghost in (True, False):
            ref = nh.Field.from_numpy(L, np.full(L.npts, 5.0), where="cpu")
            orc.assign(ref, t, use_ghost=use_ghost)
            refs[name, use_ghost] = ref.host
    return L, trees, refs


def _check_hand_path(L, trees, refs, label):
    for (name, use_ghost), want in refs.items():
        out = gpu_empty_like(L, 5.0)
        kid = nh.assign_plan(out, trees[name], use_ghost=use_ghost)
        assert kid in (_abi.KERNEL_DIVKGRAD, _abi.KERNEL_LAPL_EULER), (label, name)
        nh.assign(out, trees[name], use_ghost=use_ghost)
        assert np.array_equal(fetch(out), want), (label, name, use_ghost)


def test_hand_kernels_two_ghost_layers_every_tile():
    """With two ghost layers a ghost-inclusive <<= of a one-cell stencil chain runs
    from -1 to n+1 on every axis: the hand-written kernels then read the points
    x = -2 and y = -2 of the frame.  Every tile, through nebo_assign (the plan picks
    the hand-written kernel), equals the oracle's walk bit for bit, with the ghost
    cells assigned (use_ghost) and with the interior only."""
    L, trees, refs = _two_ghost_case((70, 37, 45), 21)
    try:
        for tile in nh.stencil_tiles():
            print(f"two ghost layers, tile {tile}", flush=True)
            nh.stencil_set_tile(0, *tile)
            nh.stencil_set_tile(1, *tile)
            _check_hand_path(L, trees, refs, tile)
    finally:
        nh.stencil_reset_tile(0)
        nh.stencil_reset_tile(1)


def test_hand_kernels_two_ghost_layers_default_tile():
    """A 254^3 interior with two ghost layers: the ghost-inclusive region is 256^3,
    for which the box-size rule picks the default tile as pair columns (128 x 8
    rows, 64-plane chunks); the interior-only region (254^3, a width that is not
    a multiple of 128) runs the default column tile."""
    n = (254, 254, 254)
    print("two ghost layers, 254^3 default tile", flush=True)
    assert nh.stencil_select_tile(0, (256, 256, 256)) == (128, 8, 64, 11)
    assert nh.stencil_select_tile(0, (254, 254, 254)) == (128, 8, 64, 7)
    L, trees, refs = _two_ghost_case(n, 22)
    _check_hand_path(L, trees, refs, "default")


def test_tile_selection_by_box_size():
    """Without a pinned or tuned tile, boxes that the default tile fills (at
    least half the CUs get a workgroup) run it (div(k grad phi) as pair columns
    where the width is a multiple of 128), smaller ones one-wave column tiles
    with short z chunks; a pin applies to every box; a tuned tile to its box
    size only."""
    nh.stencil_reset_tile(0)
    nh.stencil_reset_tile(1)
    default = (128, 8, 64, 7)
    try:
        assert nh.stencil_select_tile(0, (256, 256, 256)) == (128, 8, 64, 11)
        assert nh.stencil_select_tile(0, (250, 256, 256)) == default
        for kern in (0, 1):
            assert nh.stencil_select_tile(kern, (256, 256, 256)) == ((128, 8, 64, 11) if kern == 0 else default)
            assert nh.stencil_select_tile(kern, (512, 512, 512)) == (128, 8, 512, 7)
            assert nh.stencil_select_tile(kern, (384, 384, 384)) == (128, 8, 256, 7)
            assert nh.stencil_select_tile(kern, (128, 128, 128)) == (64, 2, 16, 3)
            assert nh.stencil_select_tile(kern, (64, 64, 64)) == (64, 2, 4, 3)
            assert nh.stencil_select_tile(kern, (16, 16, 16)) == (64, 2, 4, 3)
        nh.stencil_set_tile(0, 64, 4, 24, 3)
        assert nh.stencil_select_tile(0, (64, 64, 64)) == (64, 4, 24, 3)
        assert nh.stencil_select_tile(0, (256, 256, 256)) == (64, 4, 24, 3)
        assert nh.stencil_select_tile(1, (256, 256, 256)) == default
        n = 24
        L = nh.Layout.build((n, n, n), 1)
        rng = np.random.default_rng(2)
        k = both(L, rng.uniform(0.5, 2.0, L.npts))
        phi = both(L, rng.standard_normal(L.npts))
        c = nh.axis_coefs(nh.build_stencils(n, n, n, 1.0, 1.0, 1.0))
        best, _ = nh.stencil_autotune(gpu_empty_like(L), k, phi, c)
        assert nh.stencil_select_tile(0, (n, n, n)) == best and nh.stencil_get_tile(0) == best
        assert nh.stencil_select_tile(0, (128, 128, 128)) == (64, 2, 16, 3)   # the pin is gone
        # results do not depend on the selection
        ref = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
        orc.divkgrad(ref, k, phi, c, 0)
        r = gpu_empty_like(L)
        nh.divkgrad(r, k, phi, c, 0)
        assert np.array_equal(fetch(r), ref.host)
    finally:
        nh.stencil_reset_tile(0)
        nh.stencil_reset_tile(1)


@pytest.mark.parametrize("fn,budget", [("SIN", 1), ("COS", 1), ("TAN", 2), ("EXP", 1), ("TANH", 2),
                                       ("LOG", 1), ("LOG10", 2), ("ERF", 2), ("ERFC", 4), ("INV_ERF", 5),
                                       ("INV_ERFC", 6)])
def test_transcendental_ulp_budgets(fn, budget):
    """test/NeboTest.cpp:497-523 budgets against the oracle (host libm; erf^-1 and
    erfc^-1 restated from libm's erf/erfc, within 2 ulp of the exact values; the
    reference compares erf, erfc and the inverses with Boost.Math, which the
    correctly rounded vectors of test_special_functions_vs_correctly_rounded stand
    in for)."""
    n = (40, 30, 20)
    L = nh.Layout.build(n, 0)
    rng = np.random.default_rng(1)
    if fn in ("LOG", "LOG10"):
        x = rng.uniform(0.01, 3.0, L.npts)
    elif fn == "INV_ERF":
        x = rng.uniform(-1.0, 1.0, L.npts)
    elif fn == "INV_ERFC":
        x = rng.uniform(0.0, 2.0, L.npts)
    else:
        x = rng.uniform(-3, 3, L.npts)
    a = both(L, x)
    e = nh.Expr(fn, (nh.E(a),))
    ref = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.assign(ref, e)
    out = gpu_empty_like(L)
    nh.assign(out, e)
    got = fetch(out)
    ulp = np.abs(got - ref.host) / np.spacing(np.abs(ref.host))
    assert ulp.max() <= budget, (fn, ulp.max())


@pytest.mark.parametrize("name,fn,budget", [("erf", "ERF", 2), ("erfc", "ERFC", 4), ("erf_inv", "INV_ERF", 5),
                                            ("erfc_inv", "INV_ERFC", 6)])
def test_special_functions_vs_correctly_rounded(name, fn, budget):
    """erf, erfc, erf^-1, erfc^-1 on the GPU against correctly rounded values
    (tests/golden/special_functions.json: the reference's sin-pattern inputs of
    FieldHelper.h:82-85 plus tails and edges), within the budgets the reference
    holds the GPU to against Boost.Math (test/NeboTest.cpp:519-523); the domain
    ends give +-inf and points outside NaN."""
    x, want = orc.golden_special(name)
    L = nh.Layout.build((len(x), 1, 1), 0)
    a = both(L, x)
    out = gpu_empty_like(L)
    nh.assign(out, nh.Expr(fn, (nh.E(a),)))
    got = fetch(out)
    d = orc.ulp_distance(got, want)
    assert d.max() <= budget, (name, x[d.argmax()], got[d.argmax()], want[d.argmax()], d.max())
    if fn.startswith("INV"):
        edges = np.array([1.0, -1.0, 1.5, np.nan] if fn == "INV_ERF" else [0.0, 2.0, -0.5, 2.5])
        expect = [np.inf, -np.inf, np.nan, np.nan]
        Le = nh.Layout.build((4, 1, 1), 0)
        e = gpu_empty_like(Le)
        nh.assign(e, nh.Expr(fn, (nh.E(both(Le, edges)),)))
        assert np.array_equal(fetch(e), np.array(expect), equal_nan=True), fn


def _field_helper_values(L, start):
    """spatialops/structured/FieldHelper.h:82-85 over the whole window:
    sin(start + x + y*X + z*X*Y) with 1-based x, y, z."""
    X, Y, Z = L.glob
    z, y, x = np.meshgrid(np.arange(1, Z + 1), np.arange(1, Y + 1), np.arange(1, X + 1), indexing="ij")
    return np.sin(start + x + y * X + z * X * Y).reshape(-1)


def test_pow_two_operands_within_two_ulp():
    """pow(input1, input2) of test/NeboTest.cpp:509 with its inputs (FieldHelper.h
    sin patterns, start 0 and total): within 2 ulp of the host pow, NaN exactly
    where the host gives NaN (negative bases with non-integer exponents), as
    run_test_ignore_nan_ulp compares (NeboTest.cpp:632-637); plus integer
    exponents, zero and infinite operands."""
    L = nh.Layout.build((30, 20, 10), 1)
    total = L.npts
    a = both(L, _field_helper_values(L, 0.0))
    b = both(L, _field_helper_values(L, float(total)) * 4.0)
    out = gpu_empty_like(L)
    nh.assign(out, nh.pow_(a, b))
    ref = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.assign(ref, nh.pow_(a, b))
    got = fetch(out)
    nan = np.isnan(ref.host)
    assert np.array_equal(np.isnan(got), nan) and nan.any() and (~nan).sum() > L.npts // 3
    d = orc.ulp_distance(got[~nan], ref.host[~nan])
    assert d.max() <= 2, d.max()
    Ls = nh.Layout.build((8, 1, 1), 0)
    base = both(Ls, np.array([-2.0, -2.0, 0.0, 0.0, np.inf, 0.5, -np.inf, 1.0]))
    expo = both(Ls, np.array([3.0, -1.0, -1.0, 0.0, -2.0, np.inf, 3.0, np.nan]))
    o = gpu_empty_like(Ls)
    nh.assign(o, nh.pow_(base, expo))
    r = nh.Field.from_numpy(Ls, np.zeros(8), where="cpu")
    orc.assign(r, nh.pow_(base, expo))
    assert np.array_equal(fetch(o), r.host, equal_nan=True)


# --------------------------------------------------------------------------
# reductions, periodic fill, errors
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [(1, 1, 1), (5, 7, 3), (64, 64, 64), (258, 1, 1)])
def test_reductions(n, fused_mode):
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(sum(n))
    data = rng.standard_normal(L.npts) * 10.0
    f = both(L, data)
    for use_ghost, (mx, mn, sm, nr) in ((True, (nh.field_max, nh.field_min, nh.field_sum, nh.field_norm)),
                                        (False, (nh.field_max_interior, nh.field_min_interior,
                                                 nh.field_sum_interior, nh.field_norm_interior))):
        assert mx(f) == orc.reduce(f, _abi.RED_MAX, use_ghost)
        assert mn(f) == orc.reduce(f, _abi.RED_MIN, use_ghost)
        s_ref = orc.reduce(f, _abi.RED_SUM, use_ghost)
        arr = f.array()
        if not use_ghost:
            arr = arr[1:-1, 1:-1, 1:-1] if min(n) > 1 else arr[tuple(slice(1, -1) if g > 1 else slice(None)
                                                                   for g in L.glob[::-1])]
        exact = math.fsum(arr.ravel())
        s = sm(f)
        # the engine's double-double sum is within one rounding of the exact sum;
        # the reference's serial sum is within its own bound gamma_n * sum|x|
        assert abs(s - exact) <= 2.0 ** -52 * abs(exact) + 1e-300
        assert abs(s - s_ref) <= arr.size * 2.0 ** -53 * np.abs(arr).sum()
        nr_ref = math.sqrt(math.fsum((arr * arr).ravel()))
        assert abs(nr(f) - nr_ref) <= 4 * 2.0 ** -53 * nr_ref


def test_reduce_expression_and_nan(fused_mode):
    L = nh.Layout.build((16, 16, 16), 1)
    rng = np.random.default_rng(9)
    v = rng.standard_normal(L.npts)
    v[100] = np.nan
    f = both(L, v)
    db, ops = _ops((16, 16, 16))
    g = ops["GradX"](f)
    assert nh.field_max(f) == orc.reduce(f, _abi.RED_MAX)
    assert nh.field_min(f * 2.0 + 1.0) == orc.reduce(f * 2.0 + 1.0, _abi.RED_MIN)
    # a reduction over a stencil expression covers its valid ghosts
    w = both(L, rng.standard_normal(L.npts))
    gw = ops["GradX"](w)
    assert nh.field_max(gw) == orc.reduce(gw, _abi.RED_MAX)
    assert g is not None


def test_periodic_fill():
    for n in ((8, 8, 8), (5, 9, 3), (32, 1, 1)):
        L = nh.Layout.build(n, 1)
        rng = np.random.default_rng(1)
        d = rng.standard_normal(L.npts)
        f = both(L, d)
        nh.fill_ghosts_periodic(f)
        ref = nh.Field.from_numpy(L, d, where="cpu")
        orc.fill_periodic(ref)
        assert np.array_equal(fetch(f), ref.host)


def test_errors_on_gpu():
    L = nh.Layout.build((8, 8, 8), 1)
    a = gpu_empty_like(L)
    db, ops = _ops((8, 8, 8))
    with pytest.raises(nh.NeboError) as ei:
        nh.assign(a, a + ops["DivX"](ops["GradX"](a)))
    assert "read by a stencil" in str(ei.value)
    # operand on another (CPU) device
    h = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    s = nh.serialize(nh.E(h) + 1.0, "host")
    import ctypes as C
    d = a.desc("gpu")
    rc = nh.lib().nebo_assign(C.byref(d), s.ref, 1, None)
    assert rc == _abi.NEBO_ERR_DEVICE
    # a tree that reads no field or mask: with ghost cells its region is
    # unbounded (refused before any kernel); its interior is empty and gives the
    # fold's initial value, as the reference's loop over [0, 0) does
    c = nh.E(2.0) + 1.0
    for fn in (nh.field_max, nh.field_min, nh.field_sum, nh.field_norm):
        with pytest.raises(nh.NeboError) as ei:
            fn(c)
        assert ei.value.code == _abi.NEBO_ERR_INVALID and "unbounded" in str(ei.value)
    assert nh.field_max_interior(c) == -math.inf and nh.field_min_interior(c) == math.inf
    assert nh.field_sum_interior(c) == 0.0 and nh.field_norm_interior(c) == 0.0


def test_backend_reported_per_assignment(capfd):
    """NEBO_REPORT_BACKEND: each assignment and reduction names the kernel it ran
    (hand-written tile, compiled kernel, interpreter, masked assignment) and,
    when reporting is on, prints the reference's Starting/Finished lines."""
    n = (70, 37, 45)
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(30)
    k = both(L, rng.uniform(0.5, 2.0, L.npts))
    phi = both(L, rng.standard_normal(L.npts))
    _, ops = _ops(n)
    out = gpu_empty_like(L)
    nh.set_report_backend(True)
    try:
        nh.assign(out, _divkgrad_tree(ops, k, phi, 0))
        assert nh.last_backend().startswith("div(k grad phi) kernel, tile ")
        nh.assign(out, phi * k + 1.0)
        assert nh.last_backend().startswith("compiled kernel")
        nh.set_fused_jit(False)
        nh.assign(out, phi * k + 2.0)
        assert nh.last_backend().startswith("stack-program interpreter")
        nh.set_fused_jit(True)
        nh.field_max(phi * k)
        assert nh.last_backend().startswith("compiled reduction kernel")
        m = nh.Mask(L, [(1, 2, 3), (4, 5, 6)]).add_consumer(0)
        nh.masked_assign(m, out, 3.0)
        assert nh.last_backend() == "masked assignment (2 points)"
        torch.cuda.synchronize()
    finally:
        nh.set_report_backend(False)
        nh.set_fused_jit(True)
    printed = capfd.readouterr().out
    assert "Starting Nebo HIP div(k grad phi) kernel, tile " in printed
    assert "Finished Nebo HIP stack-program interpreter" in printed
    assert printed.count("Starting Nebo HIP") == printed.count("Finished Nebo HIP") >= 5


@pytest.mark.parametrize("n,ghosts", [((128, 37, 45), 1), ((256, 21, 9), 1), ((126, 30, 20), 2),
                                      ((128, 6, 150), 1)])
def test_pair_column_tiles_bit_exact(n, ghosts):
    """The pair-column tiles (two adjacent columns per lane, 16-byte loads and
    stores) run on boxes whose assigned width is a multiple of their 128- or
    256-column tile: each one pinned, both forms and the Euler form, rows and
    planes cut by the box end, and (two ghost layers, width 126 + 2) the
    ghost-inclusive region.  150 planes give the tiles that walk odd z chunks
    downwards several chunks, the last one short.  Every result equals the
    oracle's walk bit for bit."""
    L = nh.Layout.build(n, ghosts)
    rng = np.random.default_rng(sum(n) + ghosts)
    k = both(L, rng.uniform(0.5, 2.0, L.npts))
    phi = both(L, rng.standard_normal(L.npts))
    _, ops = _ops(n)
    refs = {}
    for form in (0, 1):
        ref = nh.Field.from_numpy(L, np.full(L.npts, 2.0), where="cpu")
        orc.assign(ref, _divkgrad_tree(ops, k, phi, form))
        refs[form] = ref.host
    pair_tiles = [t for t in nh.stencil_tiles() if t[3] >= 10]
    assert pair_tiles
    try:
        for tile in pair_tiles:
            print(f"pair tile {tile} on {n}, {ghosts} ghost layers", flush=True)
            nh.stencil_set_tile(0, *tile)
            for form in (0, 1):
                out = gpu_empty_like(L, 2.0)
                nh.assign(out, _divkgrad_tree(ops, k, phi, form))
                assert np.array_equal(fetch(out), refs[form]), (tile, form)
            if ghosts == 1:   # the Euler form (diffusion step, periodic fields, one ghost layer)
                from nebo_hip import adjoint as adj, inputs
                c = nh.axis_coefs(nh.build_stencils(*n, 1.0, 1.0, 1.0))
                pk, pp = (inputs.periodic_fill(a.host.reshape(L.glob[::-1]).copy(), 1).reshape(-1) for a in (k, phi))
                ro = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
                orc.diffusion_step(ro, nh.Field(L, host=pp), nh.Field(L, host=pk), c, 1e-4)
                o = torch.zeros_like(phi.gpu)
                adj.diffusion_step(L, o, torch.from_numpy(pp).cuda(), torch.from_numpy(pk).cuda(), c, 1e-4)
                torch.cuda.synchronize()
                assert np.array_equal(o.cpu().numpy(), ro.host), (tile, "euler")
    finally:
        nh.stencil_reset_tile(0)
