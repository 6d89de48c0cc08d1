"""The kernels the run-time compiler generates for generic trees compile for
gfx950 with the library's flags (hiprtc works without a GPU), for random trees
over every opcode, masks, mask shifts and edgeless stencils.  The GPU tests
check their results against the oracle bit for bit.

The library launches these kernels with their argument struct as the whole
argument buffer, so a kernel that reads hidden arguments (gridDim, blockDim,
global offsets) would read memory the launch never filled; every compiled
kernel's code object must therefore declare a kernarg segment of exactly the
struct's size."""
import ctypes as C
import os
import random
import re
import subprocess
import tempfile

import pytest

import nebo_hip as nh
from nebo_hip import _abi

HIPRTC = "/opt/rocm/lib/libhiprtc.so"
OPTS = [b"--offload-arch=gfx950", b"-O3", b"-ffp-contract=off", b"-Xclang", b"-target-feature", b"-Xclang",
        b"-packed-fp32-ops"]


@pytest.fixture(scope="module")
def rtc():
    if not os.path.exists(HIPRTC):
        pytest.skip("hiprtc not installed")
    return C.CDLL(HIPRTC)


READELF = "/opt/rocm/llvm/bin/llvm-readelf"
ARGS_BYTES = 1280   # sizeof(JitArgs) in csrc/nebo_jit.hip


def compile_src(rtc, src, want_code=False):
    prog = C.c_void_p()
    assert rtc.hiprtcCreateProgram(C.byref(prog), src.encode(), b"nebo_fused.hip", 0, None, None) == 0
    opts = (C.c_char_p * len(OPTS))(*OPTS)
    rc = rtc.hiprtcCompileProgram(prog, len(OPTS), opts)
    n = C.c_size_t()
    rtc.hiprtcGetProgramLogSize(prog, C.byref(n))
    log = C.create_string_buffer(n.value + 1)
    if n.value:
        rtc.hiprtcGetProgramLog(prog, log)
    code = None
    if rc == 0 and want_code:
        sz = C.c_size_t()
        rtc.hiprtcGetCodeSize(prog, C.byref(sz))
        buf = C.create_string_buffer(sz.value)
        rtc.hiprtcGetCode(prog, buf)
        code = buf.raw
    rtc.hiprtcDestroyProgram(C.byref(prog))
    if want_code:
        return rc, log.value.decode(errors="replace"), code
    return rc, log.value.decode(errors="replace")


def kernarg_sizes(code):
    """Kernarg segment size of every kernel in a code object, from its metadata note."""
    with tempfile.NamedTemporaryFile(suffix=".co") as t:
        t.write(code)
        t.flush()
        notes = subprocess.run([READELF, "--notes", t.name], capture_output=True, text=True, check=True).stdout
    return [int(v) for v in re.findall(r"\.kernarg_segment_size:\s+(\d+)", notes)]


def check_no_hidden_args(rtc, src):
    rc, log, code = compile_src(rtc, src, want_code=True)
    assert rc == 0, log + "\n" + src
    if os.path.exists(READELF):
        sizes = kernarg_sizes(code)
        assert sizes == [ARGS_BYTES], (sizes, src)


def host(L):
    f = nh.Field.empty(L, "cpu")
    f.host[:] = 0.0
    return f


def random_tree(rng, leaves, stencils, masks, depth):
    if depth == 0 or rng.random() < 0.2:
        c = rng.random()
        if c < 0.15:
            return nh.E(rng.uniform(-2, 2))
        if c < 0.25:
            return nh.E(rng.choice(masks))
        return nh.E(rng.choice(leaves))
    sub = lambda: random_tree(rng, leaves, stencils, masks, depth - 1)
    k = rng.random()
    if k < 0.3:
        return rng.choice(stencils)(sub())
    if k < 0.6:
        return nh.Expr(rng.choice(["ADD", "SUB", "MUL", "DIV", "MAX", "MIN", "POW", "EQ", "NE", "LT", "LE",
                                   "GT", "GE", "AND", "OR"]), (sub(), sub()))
    if k < 0.85:
        u = rng.choice(["NEG", "ABS", "SQRT", "SQUARE", "CUBE", "EXP", "LOG", "LOG10", "SIN", "COS", "TAN",
                        "TANH", "ERF", "ERFC", "INV_ERF", "INV_ERFC", "NOT"])
        return nh.Expr(u, (sub(),))
    if k < 0.9:
        return nh.powint(sub(), rng.choice([2, 3, -1, 5]))
    return nh.cond(sub(), sub())(sub())


def test_every_opcode_compiles(rtc):
    n = (6, 5, 4)
    L = nh.Layout.build(n, 1)
    F = nh.Layout.build(n, 1, "SSurfX")
    f, g, h = host(L), host(L), host(F)
    db = nh.build_stencils(*n, 1, 1, 1)
    m = nh.Mask(L, [(0, 0, 0)])
    stencils = [db[("Gradient", "SVol", "SSurfX")], db[("Filter", "SVol", "SVol")].sum,
                nh.Stencil([(1, 0, 0)], [2.0], kind=_abi.STENCIL_EDGELESS)]
    rng = random.Random(1)
    trees = [nh.cond(nh.mask_shift(m, (1, 0, 0)), f * g)(m, 2.0)(-g), stencils[2](f) - g,
             stencils[0](f) * h, nh.E(f), nh.E(3.0)]
    trees += [random_tree(rng, [f, g], stencils[1:], [m, nh.mask_shift(m, (0, -1, 0))], 4) for _ in range(12)]
    compiled = 0
    for t in trees:
        try:
            src = nh.fused_source(t)
        except nh.NeboError as e:   # beyond the evaluator's limits: refused before any kernel
            assert e.code == _abi.NEBO_ERR_TOO_LARGE
            continue
        check_no_hidden_args(rtc, src)
        compiled += 1
    assert compiled >= 10


@pytest.mark.parametrize("op", [_abi.RED_MAX, _abi.RED_MIN, _abi.RED_SUM, _abi.RED_NORM])
def test_reduction_kernels_compile(rtc, op):
    n = (6, 5, 4)
    L = nh.Layout.build(n, 1)
    f, g = host(L), host(L)
    db = nh.build_stencils(*n, 1, 1, 1)
    m = nh.Mask(L, [(0, 0, 0)])
    grad = db[("Gradient", "SVol", "SSurfX")]
    trees = [nh.E(f), f * g + 1.0, nh.cond(m, f)(g), grad(f) * 2.0, nh.sqrt(nh.abs_(f - g))]
    for t in trees:
        src = nh.fused_source(t, reduce_op=op)
        assert "nebo_reduce" in src
        check_no_hidden_args(rtc, src)


def test_unknown_reduction_refused():
    L = nh.Layout.build((4, 4, 4), 1)
    with pytest.raises(nh.NeboError) as ei:
        nh.fused_source(nh.E(host(L)), reduce_op=9)
    assert ei.value.code == _abi.NEBO_ERR_INVALID


def _divkgrad_tree(n=16):
    L = nh.Layout.build((n, n, n), 1)
    k, phi = host(L), host(L)
    ops = nh.basic_op_types(nh.build_stencils(n, n, n, 1.0, 1.0, 1.0))
    return sum((ops[f"Div{a}"](ops[f"InterpC2F{a}"](k) * ops[f"Grad{a}"](phi)) for a in "YZ"),
               ops["DivX"](ops["InterpC2FX"](k) * ops["GradX"](phi)))


def test_z_chain_reads_are_shared():
    """div(k grad phi) reads phi and k at 7 points each: (0,0,-1), (0,0,0), (0,0,1)
    form one z chain per field and the four x/y neighbours four chains of one.
    In a full chunk of V cells a chain of length m needs m + V - 1 planes, so the
    kernel issues 2 (3 + V - 1) + 8 V reads instead of 14 V; the tail path (a
    chunk cut by the box end) reads each cell on its own."""
    t = _divkgrad_tree()
    nh.fused_jit_column(False)
    try:
        for (block, cells, budget), V in (((0, 0, 0), 4), ((256, 4, 32), 2), ((128, 8, 96), 8),
                                          ((64, 1, 0), 1)):
            nh.fused_jit_tune(block, max_cells=cells, load_budget=budget)
            src = nh.fused_source(t)
            full = src.split("} else {")[0]
            assert full.count("const double Z") == 2 * (3 + V - 1) + 8 * V, (block, cells, budget)
            assert f"__launch_bounds__({block or 512})" in src
            if V > 1:
                tail = src.split("} else {")[1]
                assert tail.count("if (") >= V - 1 and tail.count("A.base[") == 14 * (V - 1)
    finally:
        nh.fused_jit_tune()
        nh.fused_jit_column()


def test_column_schedule_source():
    """div(k grad phi) on the column schedule with one column per lane (mode 2):
    k and phi are one family each (reads within one point in x, y and z), each
    lane holds 4 rows,
    so a family keeps rows -1 .. 4 of 3 planes plus the plane loading: 2 x 6 x 3
    loads before the z loop and 2 x 6 per step, each a raw buffer load at a
    per-lane offset from a per-plane resource.  The four x-neighbour reads per
    cell come from the adjacent lanes (4 x 4 lane exchanges per step); the values
    past the wave's edges come from one more load per family and plane for all
    rows, issued with the plane's rows and kept in the ring beside them.
    The ring of 4 planes makes the main loop 4 steps long (registers renamed, not
    copied) and a remainder loop 1 step.  Trees with masks, and pointwise trees,
    keep the flattened-plane schedule, as does every tree with the column schedule
    switched off."""
    t = _divkgrad_tree()
    try:
        nh.fused_jit_column(2)
        src = nh.fused_source(t)
    finally:
        nh.fused_jit_column()
    assert "__launch_bounds__(64)" in src and "int z = zs;   // column schedule" in src
    assert "ld2_(" not in src
    pro, loops = src.split("int z = zs;   // column schedule")
    main, rest = loops.split("for (; z < zend; ++z)")
    assert "for (; z + 3 < zend; z += 4)" in main
    ring = lambda text: len(re.findall(r"R\d+_\d+_\d+ = ld_\(rs_, o\d+_\d+\)", text))
    edge = lambda text: len(re.findall(r"V\d+_\d+ = ld_\(rs_, eo\d+_\)", text))
    assert ring(pro) == 2 * 6 * 3 and ring(main) == 4 * 2 * 6 and ring(rest) == 2 * 6
    assert main.count("__builtin_amdgcn_mov_dpp") == 4 * 16 and rest.count("__builtin_amdgcn_mov_dpp") == 16
    assert edge(pro) == 2 * 3 and edge(main) == 4 * 2 and edge(rest) == 2
    assert main.count("rdlane_(V") == 4 * 16 and rest.count("rdlane_(V") == 16
    # the main loop renames (no ring copies); the remainder loop shifts the ring
    assert not re.search(r"R\d+_\d+_\d+ = R", main) and len(re.findall(r"R\d+_\d+_\d+ = R", rest)) == 2 * 6 * 3
    # one resource per family and plane (3 planes before the loop, 2 per step),
    # and no per-read index clamp
    assert pro.count("plane_rs_(A.base[") == 6 and main.count("plane_rs_(A.base[") == 4 * 2
    assert rest.count("plane_rs_(A.base[") == 2 and "i_ < A.glo[" not in src
    n = (6, 5, 4)
    L = nh.Layout.build(n, 1)
    f, g = host(L), host(L)
    m = nh.Mask(L, [(0, 0, 0)])
    db = nh.build_stencils(*n, 1, 1, 1)
    for other in (f * g + 1.0, nh.cond(m, db[("Gradient", "SVol", "SSurfX")](f))(g)):
        assert "// column schedule" not in nh.fused_source(other)
    try:
        nh.fused_jit_column(True, 4, 16)
        src = nh.fused_source(t)
        assert "__launch_bounds__(256)" in src and "bz * 16" in src
        nh.fused_jit_column(False)
        assert "// column schedule" not in nh.fused_source(t)
    finally:
        nh.fused_jit_column()


def test_column_bounds():
    for bad in ((3, 0, 0), (-1, 0, 0), (1, 17, 0), (1, -1, 0), (1, 0, 5000), (1, 0, -3)):
        with pytest.raises(nh.NeboError) as ei:
            nh._abi.call("nebo_fused_jit_column", *bad)
        assert ei.value.code == _abi.NEBO_ERR_INVALID
    nh.fused_jit_column()


def test_tune_bounds():
    for bad in ((32, -1, 0, 0, -1), (1088, -1, 0, 0, -1), (0, -1, 3, 0, -1), (0, -1, 32, 0, -1),
                (0, -1, 0, 300, -1), (256, 96, 0, 0, -1), (256, 512, 0, 0, -1), (0, -1, 0, 0, 2)):
        with pytest.raises(nh.NeboError) as ei:
            nh.fused_jit_tune(*bad)
        assert ei.value.code == _abi.NEBO_ERR_INVALID
    nh.fused_jit_tune()


@pytest.mark.parametrize("shape,column", [((1024, 0, 8, 64, 0), (0, 0, 0)), ((64, -1, 16, 256, 1), (0, 0, 0)),
                                          ((1024, 64, 4, 48, 1), (0, 0, 0)), ((512, 128, 2, 0, 0), (0, 0, 0)),
                                          ((0, -1, 0, 0, 0), (1, 4, 1)), ((0, -1, 0, 0, 1), (1, 16, 4096))])
def test_other_shapes_compile(rtc, shape, column):
    try:
        nh.fused_jit_tune(*shape)
        nh.fused_jit_column(*column)
        n = (6, 5, 4)
        L = nh.Layout.build(n, 1)
        f, g = host(L), host(L)
        m = nh.Mask(L, [(0, 0, 0)])
        db = nh.build_stencils(*n, 1, 1, 1)
        for t in (_divkgrad_tree(), f * g + 1.0, nh.cond(m, f)(g), db[("Gradient", "SVol", "SSurfX")](f)):
            check_no_hidden_args(rtc, nh.fused_source(t))
            check_no_hidden_args(rtc, nh.fused_source(t, reduce_op=_abi.RED_SUM))
    finally:
        nh.fused_jit_tune()
        nh.fused_jit_column()


def test_column_faces_reused_across_rows_and_planes():
    """In div(k grad phi) the low y face of row r is the high y face of row r - 1,
    and the low z face of plane z the high z face of plane z - 1: the column
    kernel forms each once, carrying y faces from row to row (3 carries for 4
    rows) and z faces from plane to plane (one per row); the chunk's first
    plane takes its low faces from a prologue that forms them at plane zs - 1
    before the z loop, which has no branch on the plane.  A subtree that differs from its shifted
    partner in a constant is never taken for it.  (One column per lane here;
    test_pair_column_source covers two.)"""
    try:
        nh.fused_jit_column(2)
        src = nh.fused_source(_divkgrad_tree())
    finally:
        nh.fused_jit_column()
    assert len(set(re.findall(r"\bcy\d+_\d+", src))) == 3
    assert len(set(re.findall(r"\bcz\d+_\d+", src))) == 4
    assert "if (zq == zs)" not in src and src.count("const int z0 = zs - 1;") == 4
    n = (8, 8, 8)
    L = nh.Layout.build(n, 1)
    f = host(L)
    hi_ = nh.Stencil([(0, 0, 0), (0, 1, 0)], [1.0, 2.0])
    same = nh.Stencil([(0, -1, 0), (0, 0, 0)], [1.0, 2.0])
    other = nh.Stencil([(0, -1, 0), (0, 0, 0)], [3.0, 4.0])
    for lo_, pairs in ((same, True), (other, False)):
        try:
            nh.fused_jit_column(2)
            src = nh.fused_source(hi_(f) * 0.5 + lo_(f) * 0.25)
        finally:
            nh.fused_jit_column()
        assert "// column schedule" in src
        assert (len(set(re.findall(r"\bcy\d+_\d+", src))) > 0) == pairs


def test_pair_column_source():
    """By default div(k grad phi) runs with two adjacent columns per lane and 4
    waves stacked per block: the ring of two families at 4 rows would need 104
    doubles, so a lane holds 2 rows of both columns (72 doubles).  Every ring load and store moves 16
    bytes (ld2_ / st2_; a lane whose second column lies past the box stores its
    first alone); the x face between a lane's two cells is formed by the first
    and carried to the second, so each row takes 2 lane exchanges per family
    instead of 4.  A tree of three such families fits two columns per lane (2
    rows) only with the plane below each family's ring dropped (81 doubles
    instead of 108), which it then is."""
    src = nh.fused_source(_divkgrad_tree())
    assert "// column schedule" in src and "ncy = (A.n1 + 7) / 8" in src and "__launch_bounds__(256)" in src
    assert "const int x = A.lo0 + bx * 128 + 2 * lane_, xc_ = x;" in src
    pro, loops = src.split("int z = zs;   // column schedule")
    main, rest = loops.split("for (; z < zend; ++z)")
    assert "for (; z + 3 < zend; z += 4)" in main
    ring = lambda text: len(re.findall(r"ld2_\(rs_, o\d+_\d+, R\d+_\d+_\d+_0, R\d+_\d+_\d+_1\)", text))
    assert ring(pro) == 2 * 4 * 3 and ring(main) == 4 * 2 * 4 and ring(rest) == 2 * 4
    assert not re.search(r"R\d+_\d+_\d+ = ld_\(", src)
    assert main.count("__builtin_amdgcn_mov_dpp") == 4 * 2 * 4 and rest.count("__builtin_amdgcn_mov_dpp") == 2 * 4
    assert main.count("st2_(&A.out") == 4 * 2 and rest.count("st2_(&A.out") == 2
    assert len(set(re.findall(r"\bcx\d+_\d+", src))) == 2   # one x face per row
    assert len(set(re.findall(r"\bcy\d+_\d+_\d", src))) == 2   # one y face per column
    n = (6, 5, 4)
    L = nh.Layout.build(n, 1)
    f, g, h = host(L), host(L), host(L)
    db = nh.build_stencils(*n, 1, 1, 1)
    lap = lambda q: sum((db[("Divergence", f"SSurf{a}", "SVol")](db[("Gradient", "SVol", f"SSurf{a}")](q))
                         for a in "YZ"), db[("Divergence", "SSurfX", "SVol")](db[("Gradient", "SVol", "SSurfX")](q)))
    src3 = nh.fused_source(lap(f) + lap(g) + lap(h))
    assert "// column schedule" in src3 and "ld2_(rs_" in src3 and "ncy = (A.n1 + 7) / 8" in src3
    assert re.search(r"ld2_\(rs_, o\d+_\d+, P\d+_\d+_0, P\d+_\d+_1\)", src3)
    # one family (the Laplacian of one field) keeps one column and one wave
    src1 = nh.fused_source(lap(f))
    assert "ld2_(" not in src1 and "__launch_bounds__(64)" in src1


def test_transport_stage_drops_the_plane_below():
    """A transport stage (six fields: upwinded convective and diffusive fluxes)
    fits two columns per lane only with the ring one plane shorter: its z faces
    are carried from plane to plane, so the plane below is read only for the
    chunk's first plane, which a prologue forms from one load of that plane per
    family (P... registers) before the z loop.  div(k grad phi), which fits as
    it is, keeps the plane in its ring."""
    from nebo_hip import workloads as W
    n = 16
    L, Ls, ops, k, phi, u, dt = W.transport_problem(n)
    K = nh.Field.from_numpy(L, k, where="cpu")
    U = [nh.Field.from_numpy(l, x.reshape(-1), where="cpu") for l, x in zip(Ls, u)]
    P = [nh.Field.from_numpy(L, phi, where="cpu") for _ in range(3)]
    lo, hi = W.upwind_ops()
    s1, s2, s3 = W.rk3_stage_exprs(ops, P[0], K, U, dt, lo, hi)
    src = nh.fused_source(s2(P[1]))
    pro, loops = src.split("int z = zs;   // column schedule")
    assert "ld2_(" in src and "(zs + -1)" in pro and re.search(r"\bP\d+_\d+_0 = |ld2_\(rs_, o\d+_\d+, P\d+_", pro)
    assert "(zs + -1)" not in loops and "if (zq == zs)" not in src
    # the wave-edge values are read by every lane, then selected
    assert re.search(r"const double E0_\d+ = rdlane_\(", loops) and "? rdlane_(" not in src
    div = nh.fused_source(_divkgrad_tree())
    assert "ld2_(rs_" in div and not re.search(r"\bP\d+_\d+_\d+", div)
