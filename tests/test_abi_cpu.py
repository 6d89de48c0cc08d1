"""CPU tests of the engine library: it loads, exports every symbol the header
declares, and its host-only logic (ghost inference, tree recognition, errors)
agrees with the oracle.  No compute call touches a GPU here."""
import ctypes as C
import random

import numpy as np
import pytest

import nebo_hip as nh
from nebo_hip import _abi
import oracle_lib as orc


def test_library_exports_every_header_symbol():
    L = nh.lib()
    syms = nh.header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_device_count_without_gpu_is_zero_or_more():
    assert nh.device_count() >= 0


def host_field(layout, fill=0.0):
    f = nh.Field.empty(layout, "cpu")
    f.host[:] = fill
    return f


def _lib_limits(lhs, expr, use_ghost):
    s = nh.serialize(expr, "host")
    d = lhs.desc("host")
    lim = _abi.LimitsT()
    empty = C.c_int(0)
    rc = nh.lib().nebo_assign_limits(C.byref(d), s.ref, int(use_ghost), C.byref(lim), C.byref(empty))
    if rc:
        return ("error", rc)
    return None if empty.value else (tuple(lim.lo), tuple(lim.hi))


def _orc_limits(lhs, expr, use_ghost):
    s = nh.serialize(expr, "host")
    d = lhs.desc("host")
    lim = _abi.LimitsT()
    empty = C.c_int(0)
    rc = orc.lib().oracle_assign_limits(C.byref(d), s.ref, int(use_ghost), C.byref(lim), C.byref(empty))
    if rc:
        return ("error", rc)
    return None if empty.value else (tuple(lim.lo), tuple(lim.hi))


def random_tree(rng, leaves, stencils, masks, depth):
    r = rng.random()
    if depth == 0 or r < 0.2:
        c = rng.random()
        if c < 0.15:
            return nh.E(rng.uniform(-2, 2))
        if c < 0.25 and masks:
            return nh.E(rng.choice(masks))
        return nh.E(rng.choice(leaves))
    kind = rng.random()
    sub = lambda: random_tree(rng, leaves, stencils, masks, depth - 1)
    if kind < 0.35:
        return rng.choice(stencils)(sub())
    if kind < 0.75:
        op = rng.choice(["ADD", "SUB", "MUL", "MAX", "MIN"])
        return nh.Expr(op, (sub(), sub()))
    if kind < 0.85:
        return nh.Expr(rng.choice(["NEG", "ABS", "SQUARE"]), (sub(),))
    return nh.cond(sub() > sub(), sub())(sub())


@pytest.mark.parametrize("seed", range(40))
def test_limits_agree_with_oracle_on_random_trees(seed):
    rng = random.Random(seed)
    n = (rng.randint(1, 9), rng.randint(1, 9), rng.randint(1, 9))
    bc = (rng.random() < 0.3, rng.random() < 0.3, rng.random() < 0.3)
    g = rng.randint(0, 2)
    loc = rng.choice(["SVol", "SSurfX", "SSurfY", "XVol"])
    L = nh.Layout.build(n, g, loc, bc)
    lhs = host_field(L)
    leaves = [host_field(L), nh.Field(L.with_valid(max(g - 1, 0), g), host=host_field(L).host)]
    db = nh.build_stencils(max(n[0], 1), max(n[1], 1), max(n[2], 1), 1, 1, 1)
    stencils = [db[("Gradient", "SVol", "SSurfX")], db[("Divergence", "SSurfY", "SVol")],
                db[("Interpolant", "SVol", "XSurfY")], db[("Filter", "SVol", "SVol")].sum,
                db[("GradientZ", "SVol", "SVol")],
                nh.Stencil([(1, 0, 0)], [1.0], kind=_abi.STENCIL_EDGELESS),
                nh.Stencil([(0, -1, 0), (0, 0, 2)], [0.5, -2.0], kind=_abi.STENCIL_EDGELESS)]
    m0 = nh.Mask(L, [(0, 0, 0)])
    masks = [m0, nh.mask_shift(m0, (1, 0, -1)), nh.mask_shift(m0, (0, -2, 0))]
    for _ in range(5):
        e = random_tree(rng, leaves, stencils, masks, 4)
        for ug in (True, False):
            assert _lib_limits(lhs, e, ug) == _orc_limits(lhs, e, ug)


def test_plan_recognises_the_hot_trees():
    n = (8, 8, 8)
    L = nh.Layout.build(n, 1)
    db = nh.build_stencils(*n, 1, 1, 1)
    ops = nh.basic_op_types(db)
    k, phi, r = host_field(L), host_field(L), host_field(L)
    for where in ("host",):
        D = [ops[f"Div{a}"](ops[f"InterpC2F{a}"](k) * ops[f"Grad{a}"](phi)) for a in "XYZ"]
        assert nh.assign_plan(r, D[0] + D[1] + D[2], where=where) == _abi.KERNEL_DIVKGRAD
        Dg = [ops[f"Div{a}"](ops[f"Grad{a}"](phi) * ops[f"InterpC2F{a}"](k)) for a in "XYZ"]
        assert nh.assign_plan(r, -Dg[0] - Dg[1] - Dg[2], where=where) == _abi.KERNEL_DIVKGRAD
        # wrong axis order is not the recognised tree
        assert nh.assign_plan(r, D[1] + D[0] + D[2], where=where) == _abi.KERNEL_GENERIC
        lap = sum((ops[f"Div{a}"](ops[f"Grad{a}"](phi)) for a in "YZ"), ops["DivX"](ops["GradX"](phi)))
        assert nh.assign_plan(r, phi + 0.25 * lap, where=where) == _abi.KERNEL_LAPL_EULER
        assert nh.assign_plan(r, phi + lap * 0.25, where=where) == _abi.KERNEL_LAPL_EULER
        # in place (out is the read field) is not the specialised kernel
        assert nh.assign_plan(phi, phi + 0.25 * lap, where=where) == _abi.KERNEL_GENERIC
        assert nh.assign_plan(r, phi * k + 1.0, where=where) == _abi.KERNEL_GENERIC


def test_errors_carry_reference_prefix():
    n = (4, 4, 4)
    L = nh.Layout.build(n, 1)
    f = host_field(L)
    bad = _abi.NodeT()
    bad.op = 99
    arr = (_abi.NodeT * 1)(bad)
    e = _abi.ExprT(1, arr, 0, None, 0, None, 0, None)
    d = f.desc("host")
    rc = nh.lib().nebo_assign_limits(C.byref(d), C.byref(e), 1, C.byref(_abi.LimitsT()), C.byref(C.c_int()))
    assert rc == _abi.NEBO_ERR_INVALID
    assert nh.lib().nebo_last_error().decode().startswith("Nebo error in Nebo Assignment:")


def test_cpu_lhs_is_refused_loudly():
    """The engine evaluates on GPUs only; a CPU-resident LHS is an error, never a
    silent CPU evaluation."""
    L = nh.Layout.build((4, 4, 4), 1)
    f, g = host_field(L), host_field(L)
    s = nh.serialize(g * 2.0, "host")
    d = f.desc("host")
    rc = nh.lib().nebo_assign(C.byref(d), s.ref, 1, None)
    assert rc == _abi.NEBO_ERR_DEVICE
    assert "GPU only" in nh.lib().nebo_last_error().decode()


def test_stencil_on_lhs_is_refused():
    L = nh.Layout.build((4, 4, 4), 1)
    f = host_field(L)
    f.device = 0
    db = nh.build_stencils(4, 4, 4, 1, 1, 1)
    gx = db[("GradientX", "SVol", "SVol")]
    s = nh.serialize(f + gx(f), "gpu" if False else "host")
    d = f.layout.desc(f.host.ctypes.data, 0)
    # make every field "GPU 0" so the device check passes and the aliasing check decides
    for i in range(s.expr.n_fields):
        s.expr.fields[i].device = 0
    rc = nh.lib().nebo_assign(C.byref(d), s.ref, 1, None)
    assert rc == _abi.NEBO_ERR_INVALID
    assert "read by a stencil" in nh.lib().nebo_last_error().decode()


def test_pending_entry_points_fail_loudly():
    out = C.c_void_p()
    rc = nh.lib().nebo_batch_create(None, 0, None, 0.0, None, C.byref(out))
    assert rc in (_abi.NEBO_ERR_UNSUPPORTED, _abi.NEBO_ERR_INVALID)


def test_batch_tile_checked_on_entry():
    """Column tiles (x = 0) take 1, 2 or 4 rows per lane and at least one plane;
    the tile is checked before any device call."""
    arr = (_abi.PatchT * 1)()
    arr[0].a, arr[0].b = 16, 32
    arr[0].n[:] = [4, 4, 4]
    c = nh.axis_coefs(nh.build_stencils(4, 4, 4, 1, 1, 1))
    for bad in ((0, 3, 8), (0, 2, 0), (0, 0, 4), (96, 0, 4), (64, 32, 4)):
        out = C.c_void_p()
        t = (C.c_int32 * 3)(*bad)
        assert nh.lib().nebo_batch_create(arr, 1, C.byref(c), 0.1, t, C.byref(out)) == _abi.NEBO_ERR_INVALID, bad
        assert not out.value


def test_backward_checkpoint_interval_checked_on_entry():
    """An interval above 1024 steps is refused before any work is queued (no
    copy, no clearing of g_k): the descriptors here are host memory marked as
    GPU 0, so reaching a HIP call would fail differently."""
    L = nh.Layout.build((4, 4, 4), 1)
    fs = [host_field(L, 1.0) for _ in range(3)]
    ds = []
    for f in fs:
        d = f.layout.desc(f.host.ctypes.data, 0)
        ds.append(d)
    gk = np.full(L.npts, 3.0)
    work = np.zeros(L.npts)
    c = nh.axis_coefs(nh.build_stencils(4, 4, 4, 1, 1, 1))
    rc = nh.lib().nebo_diffusion_backward(C.byref(ds[0]), gk.ctypes.data_as(C.c_void_p), C.byref(ds[1]),
                                          C.byref(ds[2]), C.byref(c), 1e-3, 2048, 2048,
                                          work.ctypes.data_as(C.c_void_p), work.ctypes.data_as(C.c_void_p), None)
    assert rc == _abi.NEBO_ERR_TOO_LARGE
    assert "checkpoint interval" in nh.lib().nebo_last_error().decode()
    assert np.all(gk == 3.0)


def test_backend_reporting_switch():
    """nebo_set_report_backend / nebo_last_backend exist and the switch is
    accepted both ways (the names themselves are checked on the GPU)."""
    lib = nh.lib()
    assert lib.nebo_set_report_backend(1) == 0
    assert lib.nebo_set_report_backend(0) == 0
    assert isinstance(lib.nebo_last_backend(), bytes)
