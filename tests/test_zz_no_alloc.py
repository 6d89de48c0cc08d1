"""nebo_reduce_async_noalloc (the reduction entry point that allocates nothing)
against nebo_reduce_async, through the C-ABI: two calls in a row, the second on
smaller fields, every reduction op with and without ghosts, on trees the
allocating call compiled first (same kernel: equal bit for bit) and on a tree
shape seen first by the allocation-free call (interpreter kernels: max/min
equal, sum/norm within one rounding of the compiled kernel's result)."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

import nebo_hip as nh
from nebo_hip import _abi

pytestmark = pytest.mark.gpu

OPS = (_abi.RED_MAX, _abi.RED_MIN, _abi.RED_SUM, _abi.RED_NORM)


def _reduce(fn, tree, op, use_ghost, out, stream):
    s = nh.serialize(tree, "gpu")
    rc = getattr(nh.lib(), fn)(s.ref, op, use_ghost, C.c_void_p(out.data_ptr()), C.c_void_p(stream.cuda_stream))
    _abi.check(rc)
    return s


def _fields(n, seed):
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(seed)
    k = nh.Field.from_numpy(L, rng.uniform(0.5, 2.0, L.npts), where="gpu")
    phi = nh.Field.from_numpy(L, rng.standard_normal(L.npts), where="gpu")
    return L, k, phi


def _close(a, b, op, scale):
    if op in (_abi.RED_MAX, _abi.RED_MIN):
        return a == b
    return abs(a - b) <= 2.0 ** -51 * abs(b) + 2.0 ** -90 * scale


def test_noalloc_reduction_matches_allocating_entry_point():
    stream = torch.cuda.Stream()
    for call, n in enumerate([(96, 80, 72), (40, 30, 20)]):
        print(f"no-alloc reductions, call {call}, box {n}", flush=True)
        L, k, phi = _fields(n, 40 + call)
        ops = nh.basic_op_types(nh.build_stencils(*n, 1.0, 1.0, 1.0))
        trees = [nh.E(phi), phi * k + 1.0, ops["DivX"](ops["GradX"](phi)) * k]
        got = torch.zeros(2 * len(trees) * len(OPS) * 2, dtype=torch.float64, device="cuda")
        slot = 0
        keep = []
        with torch.cuda.stream(stream):
            for t in trees:
                for op in OPS:
                    for ug in (1, 0):
                        keep.append(_reduce("nebo_reduce_async", t, op, ug, got[slot], stream))
                        keep.append(_reduce("nebo_reduce_async_noalloc", t, op, ug, got[slot + 1], stream))
                        slot += 2
            stream.synchronize()
        v = got.cpu().numpy()
        for i in range(0, slot, 2):
            assert v[i] == v[i + 1] or (math.isnan(v[i]) and math.isnan(v[i + 1])), (call, i, v[i], v[i + 1])


def test_noalloc_interpreter_path_and_repeat():
    """A tree shape first met by the allocation-free call runs on the
    interpreter's kernels; the allocating call then compiles it.  Both orders,
    twice in a row (the scratch keeps the previous call's partials), the second
    time on a smaller box."""
    stream = torch.cuda.Stream()
    for call, n in enumerate([(64, 50, 33), (20, 17, 9)]):
        print(f"no-alloc interpreter path, call {call}, box {n}", flush=True)
        L, k, phi = _fields(n, 60 + call)
        tree = nh.sqrt(nh.abs_(phi * (k - 0.25 * call))) * (2.0 + call) + nh.max_(phi, k)
        out = torch.zeros(2 * len(OPS), dtype=torch.float64, device="cuda")
        keep = []
        with torch.cuda.stream(stream):
            for j, op in enumerate(OPS):
                keep.append(_reduce("nebo_reduce_async_noalloc", tree, op, 1, out[2 * j], stream))
                keep.append(_reduce("nebo_reduce_async", tree, op, 1, out[2 * j + 1], stream))
            stream.synchronize()
        v = out.cpu().numpy()
        scale = float(torch.abs(phi.gpu).max().item() + torch.abs(k.gpu).max().item()) * L.npts
        for j, op in enumerate(OPS):
            assert _close(v[2 * j], v[2 * j + 1], op, scale), (call, op, v[2 * j], v[2 * j + 1])
