"""Batched evaluation of many uneven patches (nebo_batch_* of include/nebo_hip.h)."""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from .api import _stream


class Batch:
    """Patches with two state blocks each; one Laplacian + Euler step per call,
    evaluated by one persistent launch over a device patch table."""

    def __init__(self, blocks_a, blocks_b, dims, coefs, s, tile=None):
        n = len(dims)
        arr = (_abi.PatchT * n)()
        for i, (a, b, d) in enumerate(zip(blocks_a, blocks_b, dims)):
            assert a.numel() == (d[0] + 2) * (d[1] + 2) * (d[2] + 2) == b.numel()
            arr[i].a = a.data_ptr()
            arr[i].b = b.data_ptr()
            arr[i].n[:] = list(d)
        self._keep = (blocks_a, blocks_b)
        t = (C.c_int32 * 3)(*tile) if tile is not None else None
        h = C.c_void_p()
        _abi.call("nebo_batch_create", arr, n, C.byref(coefs), float(s), t, C.byref(h))
        self.h = h

    def step(self, swap=0, stream=None):
        _abi.call("nebo_batch_step", self.h, int(swap), _stream(stream))

    def step_per_patch(self, swap=0, stream=None):
        _abi.call("nebo_batch_step_per_patch", self.h, int(swap), _stream(stream))

    def graph_build(self, steps, stream=None):
        _abi.call("nebo_batch_graph_build", self.h, int(steps), _stream(stream))

    def graph_launch(self, stream=None):
        _abi.call("nebo_batch_graph_launch", self.h, _stream(stream))

    def autotune(self, stream=None):
        t = (C.c_int32 * 3)()
        _abi.call("nebo_batch_autotune", self.h, t, _stream(stream))
        return tuple(t)

    def info(self):
        t = (C.c_int32 * 3)()
        ni, nb = C.c_int32(), C.c_int32()
        _abi.call("nebo_batch_info", self.h, t, C.byref(ni), C.byref(nb))
        return tuple(t), ni.value, nb.value

    def close(self):
        if self.h:
            _abi.lib().nebo_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def alloc_patches(dims, device=0):
    """Two zero-filled (n+2)^3 blocks per patch."""
    a = [torch.zeros((d[0] + 2) * (d[1] + 2) * (d[2] + 2), dtype=torch.float64, device=f"cuda:{device}")
         for d in dims]
    b = [torch.zeros_like(x) for x in a]
    return a, b
