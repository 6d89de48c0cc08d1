"""Repeated calls through the C-ABI with descriptors identical to an earlier
call replay the launches that call planned (the library's call memo).  These
tests hold the replays to what a freshly planned call gives: results equal to
the oracle's walk on the current input values, the backend each call reports,
the result slot each reduction names, the stream each call names, and every
setting that changes planning (a pinned tile, the run-time compiler switch)
taking effect on the next call."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

import nebo_hip as nh
from nebo_hip import _abi
from nebo_hip.api import _stream
from nebo_hip.expr import serialize, E
import oracle_lib as orc

pytestmark = pytest.mark.gpu


def both(layout, data):
    return nh.Field.from_numpy(layout, data, where="both")


def fetch(f):
    torch.cuda.synchronize()
    return f.gpu.cpu().numpy()


def scale(f, a):
    """Multiply a field's values on the GPU and in its host copy alike."""
    f.gpu.mul_(a)
    f.host *= a


def _trees(n, k, phi):
    db = nh.build_stencils(*n, 1.0, 1.0, 1.0)
    ops = nh.basic_op_types(db)
    div = sum((ops[f"Div{a}"](ops[f"InterpC2F{a}"](k) * ops[f"Grad{a}"](phi)) for a in "YZ"),
              ops["DivX"](ops["InterpC2FX"](k) * ops["GradX"](phi)))
    lap = ops["DivX"](ops["GradX"](phi)) + ops["DivY"](ops["GradY"](phi)) + ops["DivZ"](ops["GradZ"](phi))
    return [("div(k grad phi), hand-written", div, False, True),
            ("div(k grad phi), compiled column kernel", div, True, True),
            ("Laplacian + Euler, hand-written", phi + 0.01 * lap, False, True),
            ("pointwise, compiled", phi * k + phi / k, False, True),
            ("pointwise, interpreter", nh.cond(phi > 0.0, phi * k)(phi - k), False, False)]


@pytest.mark.parametrize("case", range(5))
def test_replayed_assignments_follow_their_inputs(case):
    """Three identical calls into one output, the input scaled between them:
    every call (the first planned, the others replayed) equals the oracle on
    the values of its time, and reports the same backend."""
    n = (130, 37, 29)
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(50 + case)
    k = both(L, rng.uniform(0.5, 2.0, L.npts))
    phi = both(L, rng.standard_normal(L.npts))
    name, tree, generic, jit = _trees(n, k, phi)[case]
    print(f"memo assignment: {name}", flush=True)
    out = nh.Field.empty(L, "gpu")
    out.gpu.fill_(7.0)
    nh.set_fused_jit(jit)
    try:
        backends = []
        for step in range(3):
            nh.assign(out, tree, generic=generic)
            backends.append(nh.last_backend())
            ref = nh.Field.from_numpy(L, np.full(L.npts, 7.0), where="cpu")
            orc.assign(ref, tree)
            got = fetch(out)
            # the oracle writes the region the assignment covers; the rest keeps 7.0
            assert np.array_equal(got, ref.host), f"{name}, call {step}"
            scale(phi, 1.5)
    finally:
        nh.set_fused_jit(True)
    assert backends[0] == backends[1] == backends[2], backends
    assert ("interpreter" in backends[0]) == (not jit)


def test_replayed_reductions_write_their_own_slots():
    """One reduction tree reduced into six result slots of a device array
    (asynchronous and allocation-free entry points, max and sum), the input
    scaled between the calls: each slot holds the reduction of the values of
    its call."""
    n = (66, 40, 31)
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(60)
    phi = both(L, rng.standard_normal(L.npts))
    k = both(L, rng.uniform(0.5, 2.0, L.npts))
    lib = _abi.lib()
    res = torch.zeros(24, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    sp = _stream(stream)
    want = []
    for step in range(6):
        ef = serialize(E(phi * k), "gpu")
        base = res.data_ptr() + 32 * step
        for j, (fn, op) in enumerate((("nebo_reduce_async", _abi.RED_MAX), ("nebo_reduce_async", _abi.RED_SUM),
                                      ("nebo_reduce_async_noalloc", _abi.RED_MAX),
                                      ("nebo_reduce_async_noalloc", _abi.RED_SUM))):
            _abi.check(getattr(lib, fn)(ef.ref, op, 1, C.c_void_p(base + 8 * j), sp))
        # max as the oracle's walk finds it; the engine's double-double sum is
        # within one rounding of the exact sum of the products
        want.append((orc.reduce(phi * k, _abi.RED_MAX), math.fsum((phi.array() * k.array()).ravel())))
        stream.synchronize()
        scale(phi, -1.25)
    got = res.cpu().numpy().reshape(6, 4)
    for step, (mx, sm) in enumerate(want):
        assert got[step, 0] == mx and got[step, 2] == mx, step
        for j in (1, 3):
            assert abs(got[step, j] - sm) <= 2.0 ** -52 * abs(sm), (step, j)


def test_planning_settings_take_effect_on_the_next_call():
    """A pinned tile, its removal and the run-time compiler switch each change
    what the next identical call launches; results stay equal to the oracle."""
    n = (128, 37, 45)
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(61)
    k = both(L, rng.uniform(0.5, 2.0, L.npts))
    phi = both(L, rng.standard_normal(L.npts))
    tree = _trees(n, k, phi)[0][1]
    ref = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.assign(ref, tree)
    out = nh.Field.empty(L, "gpu")
    seen = []
    try:
        for tile in (None, (64, 4, 16, 0), (128, 8, 64, 11), None):
            print(f"memo tile {tile}", flush=True)
            if tile is None:
                nh.stencil_reset_tile(0)
            else:
                nh.stencil_set_tile(0, *tile)
            for _ in range(2):
                out.gpu.zero_()
                nh.assign(out, tree)
                assert np.array_equal(fetch(out), ref.host), tile
            seen.append(nh.last_backend())
            if tile is not None:
                assert f"tile {tile[0]}x{tile[1]}x{tile[2]}" in seen[-1], seen[-1]
    finally:
        nh.stencil_reset_tile(0)
    assert seen[0] == seen[3] and seen[1] != seen[0]
    pointwise = phi * k + 1.0
    for jit in (True, False, True):
        nh.set_fused_jit(jit)
        try:
            nh.assign(out, pointwise)
            assert ("interpreter" in nh.last_backend()) == (not jit)
        finally:
            nh.set_fused_jit(True)
    ref2 = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.assign(ref2, pointwise)
    assert np.array_equal(fetch(out), ref2.host)


def test_replays_run_on_the_stream_they_name(capfd):
    """An identical call on another stream runs there (ordered after that
    stream's earlier work), and a replayed call still prints the reference's
    Starting/Finished lines when reporting is on."""
    n = (64, 20, 24)
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(62)
    k = both(L, rng.uniform(0.5, 2.0, L.npts))
    phi = both(L, rng.standard_normal(L.npts))
    tree = phi * k + 2.0
    out = nh.Field.empty(L, "gpu")
    nh.assign(out, tree)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    nh.set_report_backend(True)
    try:
        with torch.cuda.stream(side):
            phi.gpu.mul_(3.0)          # queued on the side stream before the assignment
            nh.assign(out, tree, stream=side)
        side.synchronize()
    finally:
        nh.set_report_backend(False)
    phi.host *= 3.0
    ref = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.assign(ref, tree)
    assert np.array_equal(fetch(out), ref.host)
    printed = capfd.readouterr().out
    assert printed.count("Starting Nebo HIP compiled kernel") == 1
    assert printed.count("Finished Nebo HIP compiled kernel") == 1


def test_failed_calls_are_not_recorded():
    """A call that fails is not recorded: the identical call fails again with the
    same error (an in-place read by a stencil, an unbounded reduction), and a
    valid call on the same fields afterwards runs as a freshly planned one."""
    n = (24, 20, 16)
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(63)
    a = both(L, rng.standard_normal(L.npts))
    ops = nh.basic_op_types(nh.build_stencils(*n, 1.0, 1.0, 1.0))
    bad = a + ops["DivX"](ops["GradX"](a))
    for _ in range(3):
        with pytest.raises(nh.NeboError) as ei:
            nh.assign(a, bad)
        assert "read by a stencil" in str(ei.value)
    c = nh.E(2.0) + 1.0
    for _ in range(3):
        with pytest.raises(nh.NeboError) as ei:
            nh.field_max(c)
        assert ei.value.code == _abi.NEBO_ERR_INVALID and "unbounded" in str(ei.value)
    out = nh.Field.empty(L, "gpu")
    good = ops["DivX"](ops["GradX"](a))
    for _ in range(2):
        out.gpu.zero_()
        nh.assign(out, good)
        ref = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
        orc.assign(ref, good)
        assert np.array_equal(fetch(out), ref.host)


def test_concurrent_threads_replay_their_own_calls():
    """Two host threads, each on its own stream, repeat their own assignment and
    reduction 200 times while the other does the same (the memo is shared, the
    recording is per thread): every result equals the oracle's."""
    import threading
    n = (48, 40, 36)
    L = nh.Layout.build(n, 1)
    ops = nh.basic_op_types(nh.build_stencils(*n, 1.0, 1.0, 1.0))
    lib = _abi.lib()
    jobs = []
    for j in range(2):
        rng = np.random.default_rng(70 + j)
        k = both(L, rng.uniform(0.5, 2.0, L.npts))
        phi = both(L, rng.standard_normal(L.npts))
        tree = ops["DivX"](ops["InterpC2FX"](k) * ops["GradX"](phi)) + ops["DivY"](ops["InterpC2FY"](k) * ops["GradY"](phi)) \
            + ops["DivZ"](ops["InterpC2FZ"](k) * ops["GradZ"](phi))
        tree = tree if j == 0 else phi * k + 3.0
        out = nh.Field.empty(L, "gpu")
        out.gpu.zero_()
        res = torch.zeros(1, dtype=torch.float64, device="cuda")
        jobs.append(dict(tree=tree, out=out, desc=out.desc("gpu"), s=serialize(E(tree), "gpu"),
                         red=serialize(E(phi * k), "gpu"), res=res, stream=torch.cuda.Stream(), err=[],
                         k=k, phi=phi))
    torch.cuda.synchronize()

    def run(job):
        try:
            sp = _stream(job["stream"])
            for _ in range(200):
                _abi.check(lib.nebo_assign(C.byref(job["desc"]), job["s"].ref, 1, sp))
                _abi.check(lib.nebo_reduce_async(job["red"].ref, _abi.RED_MAX, 1,
                                                 C.c_void_p(job["res"].data_ptr()), sp))
            job["stream"].synchronize()
        except Exception as ex:   # reported by the main thread
            job["err"].append(repr(ex))
    threads = [threading.Thread(target=run, args=(job,)) for job in jobs]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    for j, job in enumerate(jobs):
        assert not job["err"], job["err"]
        ref = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
        orc.assign(ref, job["tree"])
        assert np.array_equal(fetch(job["out"]), ref.host), j
        assert job["res"].item() == orc.reduce(job["phi"] * job["k"], _abi.RED_MAX), j
