"""The kernels the run-time compiler generates for generic trees compile for
gfx950 with the library's flags (hiprtc works without a GPU), for random trees
over every opcode, masks, mask shifts and edgeless stencils.  The GPU tests
check their results against the oracle bit for bit."""
import ctypes as C
import os
import random

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


def compile_src(rtc, src):
    prog = C.c_void_p()
    assert rtc.hiprtcCreateProgram(C.byref(prog), src.encode(), b"nebo_fused.hip", 0, None, None) == 0
    opts = (C.c_char_p * len(OPTS))(*OPTS)
    rc = rtc.hiprtcCompileProgram(prog, len(OPTS), opts)
    n = C.c_size_t()
    rtc.hiprtcGetProgramLogSize(prog, C.byref(n))
    log = C.create_string_buffer(n.value + 1)
    if n.value:
        rtc.hiprtcGetProgramLog(prog, log)
    rtc.hiprtcDestroyProgram(C.byref(prog))
    return rc, log.value.decode(errors="replace")


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
                        "TANH", "ERF", "ERFC", "NOT"])
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
        rc, log = compile_src(rtc, src)
        assert rc == 0, log + "\n" + src
        compiled += 1
    assert compiled >= 10
