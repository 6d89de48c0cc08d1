"""Assignment, reductions and the direct kernel entry points of the binding.

``assign(lhs, rhs)`` is ``lhs <<= rhs`` (NeboAssignment.h:29-133),
``interior_assign`` is its useGhost=false variant (NeboAssignment.h:135-244),
``masked_assign`` writes only the mask's points (NeboAssignment.h:246-313) and the
``field_*`` functions are the reductions of NeboReductions.h:28-240.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from .expr import E, serialize
from .fields import Field, Mask


def _stream(stream=None):
    if stream is None:
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, torch.cuda.Stream):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(stream)


def assign(lhs: Field, rhs, use_ghost=True, stream=None, generic=False):
    s = serialize(E(rhs), "gpu")
    d = lhs.desc("gpu")
    fn = "nebo_assign_generic" if generic else "nebo_assign"
    _abi.call(fn, C.byref(d), s.ref, int(use_ghost), _stream(stream))
    return lhs


def interior_assign(lhs: Field, rhs, stream=None, generic=False):
    return assign(lhs, rhs, use_ghost=False, stream=stream, generic=generic)


def assign_limits(lhs: Field, rhs, use_ghost=True, where="gpu"):
    """Loop limits [lo, hi) per axis that lhs <<= rhs covers, or None if empty."""
    s = serialize(E(rhs), where)
    d = lhs.desc(where)
    lim = _abi.LimitsT()
    empty = C.c_int(0)
    _abi.call("nebo_assign_limits", C.byref(d), s.ref, int(use_ghost), C.byref(lim), C.byref(empty))
    if empty.value:
        return None
    return tuple(lim.lo), tuple(lim.hi)


def assign_plan(lhs: Field, rhs, use_ghost=True, where="gpu") -> int:
    s = serialize(E(rhs), where)
    d = lhs.desc(where)
    k = C.c_int(-1)
    _abi.call("nebo_assign_plan", C.byref(d), s.ref, int(use_ghost), C.byref(k))
    return k.value


def masked_assign(mask: Mask, lhs: Field, rhs, stream=None):
    return assign_points(mask.points, lhs, rhs, stream)


def assign_points(points, lhs: Field, rhs, stream=None):
    """lhs(p) = rhs(p) for each interior-relative (x, y, z) point p of the list."""
    import numpy as np
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.int32).reshape(-1, 3))
    s = serialize(E(rhs), "gpu")
    d = lhs.desc("gpu")
    _abi.call("nebo_masked_assign", C.byref(d), C.c_void_p(pts.ctypes.data),
              int(len(pts)), s.ref, _stream(stream))
    return lhs


def _reduce(expr, op, use_ghost, stream=None):
    s = serialize(E(expr), "gpu")
    out = C.c_double(0.0)
    _abi.call("nebo_reduce", s.ref, op, int(use_ghost), C.byref(out), _stream(stream))
    return out.value


def field_max(e, stream=None): return _reduce(e, _abi.RED_MAX, True, stream)
def field_min(e, stream=None): return _reduce(e, _abi.RED_MIN, True, stream)
def field_sum(e, stream=None): return _reduce(e, _abi.RED_SUM, True, stream)
def field_norm(e, stream=None): return _reduce(e, _abi.RED_NORM, True, stream)
def field_max_interior(e, stream=None): return _reduce(e, _abi.RED_MAX, False, stream)
def field_min_interior(e, stream=None): return _reduce(e, _abi.RED_MIN, False, stream)
def field_sum_interior(e, stream=None): return _reduce(e, _abi.RED_SUM, False, stream)
def field_norm_interior(e, stream=None): return _reduce(e, _abi.RED_NORM, False, stream)


def fill_ghosts_periodic(f: Field, stream=None):
    d = f.desc("gpu")
    _abi.call("nebo_fill_ghosts_periodic", C.byref(d), _stream(stream))
    return f


def divkgrad(r: Field, k: Field, phi: Field, coefs, form=0, stream=None):
    dr, dk, dp = r.desc("gpu"), k.desc("gpu"), phi.desc("gpu")
    _abi.call("nebo_divkgrad", C.byref(dr), C.byref(dk), C.byref(dp), C.byref(coefs), int(form),
              _stream(stream))
    return r


def lapl_euler(out: Field, phi: Field, coefs, s, stream=None):
    do, dp = out.desc("gpu"), phi.desc("gpu")
    _abi.call("nebo_lapl_euler", C.byref(do), C.byref(dp), C.byref(coefs), float(s), _stream(stream))
    return out


def set_report_backend(on: bool):
    """Print "Starting Nebo HIP <kernel>" / "Finished Nebo HIP <kernel>" around
    every assignment and reduction (the reference's NEBO_REPORT_BACKEND; the
    NEBO_REPORT_BACKEND environment variable switches it on too)."""
    _abi.call("nebo_set_report_backend", int(bool(on)))


def last_backend() -> str:
    """The kernel the calling thread's last assignment or reduction ran."""
    return _abi.lib().nebo_last_backend().decode()


def set_fused_jit(enable: bool):
    """Run generic trees as kernels compiled at run time (True, the default) or
    on the stack-program interpreter (False); both give the same bits."""
    _abi.call("nebo_set_fused_jit", int(bool(enable)))


def fused_jit_stats():
    """(kernels compiled, cache hits, failed compiles) of the run-time compiler."""
    v = [C.c_int32() for _ in range(3)]
    _abi.call("nebo_fused_jit_stats", *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def fused_jit_column(enable=True, waves=0, zchunk=0):
    """Column schedule of the compiled assignment kernels (on by default): for
    trees without masks whose reads of each field stay within one point in x, y
    and z, a lane holds one column (or two adjacent ones where the register ring
    allows), each lane 1-4 rows and `zchunk` planes, and each plane of each field
    is loaded once.  enable: True/1 on, 2 on with one column per lane, False/0
    off.  0 keeps the defaults (1 wave per block; planes per block from the box
    size: 128, 64, 32, 16 or 8)."""
    mode = 2 if enable == 2 else int(bool(enable))
    _abi.call("nebo_fused_jit_column", mode, int(waves), int(zchunk))


def fused_jit_tune(block_threads=0, tile_x=-1, max_cells=0, load_budget=0, xcd_remap=-1):
    """Launch shape of the compiled assignment kernels; the defaults restore the
    library's own (512 threads over the flattened plane, up to 4 z cells per
    thread within 48 distinct reads, XCD renumbering on)."""
    _abi.call("nebo_fused_jit_tune", int(block_threads), int(tile_x), int(max_cells), int(load_budget),
              int(xcd_remap))


def fused_source(rhs, where="host", reduce_op=None) -> str:
    """The kernel source the run-time compiler generates for rhs: the
    assignment kernel, or with reduce_op (RED_MAX, RED_MIN, RED_SUM, RED_NORM)
    the reduction kernel."""
    s = serialize(E(rhs), where)
    op = -1 if reduce_op is None else int(reduce_op)
    need = C.c_int32()
    _abi.call("nebo_fused_source", s.ref, op, None, 0, C.byref(need))
    buf = C.create_string_buffer(need.value)
    _abi.call("nebo_fused_source", s.ref, op, buf, need.value, C.byref(need))
    return buf.value.decode()


def device_count() -> int:
    n = C.c_int(0)
    _abi.call("nebo_device_count", C.byref(n))
    return n.value


KERNEL_DIVKGRAD, KERNEL_LAPL = 0, 1


def stencil_tiles():
    """(tx, ty, zchunk, variant) tiles instantiated for the hand-written stencil
    kernels; variant 0 = direct loads, 1/2 = LDS-staged rows with 1/2 planes of
    prefetch, 3 = column schedule (tx/64 waves in x, ty rows per lane), 4 = column
    schedule computing each x face once, 5/6 = ty/4 resp. ty/2 column waves of 4
    resp. 2 rows stacked in y, 7/8 = variant 5 with loads 3 resp. 4 planes ahead."""
    n = C.c_int32()
    _abi.call("nebo_stencil_tiles", C.byref(n), None, None, None, None)
    tx, ty, zc, var = ((C.c_int32 * n.value)() for _ in range(4))
    _abi.call("nebo_stencil_tiles", C.byref(n), tx, ty, zc, var)
    return list(zip(tx, ty, zc, var))


def stencil_set_tile(kernel, tx, ty, zchunk, variant=0):
    _abi.call("nebo_stencil_set_tile", int(kernel), int(tx), int(ty), int(zchunk), int(variant))


def stencil_get_tile(kernel=0):
    """The tile last pinned or tuned for kernel (the default tile if neither)."""
    v = [C.c_int32() for _ in range(4)]
    _abi.call("nebo_stencil_get_tile", int(kernel), *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def stencil_reset_tile(kernel=0):
    """Remove a pinned tile: tuned tiles and the box-size rule apply again."""
    _abi.call("nebo_stencil_set_tile", int(kernel), 0, 0, 0, 0)


def stencil_select_tile(kernel, box, device=0):
    """The tile a launch of kernel over a box of interior size box (nx, ny, nz)
    on device would use."""
    v = [C.c_int32() for _ in range(4)]
    _abi.call("nebo_stencil_select_tile", int(kernel), int(device), *[int(b) for b in box],
              *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def stencil_autotune(r: Field, k, phi: Field, coefs, stream=None, kernel=KERNEL_DIVKGRAD):
    """Time every tile of one stencil kernel (0: div(k grad phi) into r; 1:
    phi + s*lapl(phi) into r, k unused) and keep the fastest.  Returns
    ((tx, ty, zchunk, variant), {tile: ms})."""
    tiles = stencil_tiles()
    times = (C.c_float * len(tiles))()
    v = [C.c_int32() for _ in range(4)]
    dr, dp = r.desc("gpu"), phi.desc("gpu")
    dk = C.byref(k.desc("gpu")) if k is not None else None
    _abi.call("nebo_stencil_autotune", int(kernel), C.byref(dr), dk, C.byref(dp), C.byref(coefs),
              _stream(stream), *[C.byref(x) for x in v], times)
    return tuple(x.value for x in v), dict(zip(tiles, list(times)))


def lower_stats(rhs, where="host"):
    """(instructions, stack depth, distinct reads) of the fused program for rhs."""
    s = serialize(E(rhs), where)
    n, d, l = C.c_int32(), C.c_int32(), C.c_int32()
    _abi.call("nebo_lower_stats", s.ref, C.byref(n), C.byref(d), C.byref(l))
    return n.value, d.value, l.value
