"""ctypes view of include/nebo_hip.h and the loader of lib/libnebo_hip.so.

The library is the engine; this module only mirrors its C structures and turns a
non-zero status into :class:`NeboError` carrying the library's message (which
starts with "Nebo error in <context>:", as the reference's exceptions do,
spatialops/NeboLhs.h:107-153).
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (load torch's HIP runtime first so both share one runtime)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "..", "lib", "libnebo_hip.so"))

NEBO_OK = 0
NEBO_ERR_INVALID = 1
NEBO_ERR_GHOST = 2
NEBO_ERR_DEVICE = 3
NEBO_ERR_HIP = 4
NEBO_ERR_TOO_LARGE = 5
NEBO_ERR_UNSUPPORTED = 6
NEBO_CPU_INDEX = -1
NEBO_GHOST_MAX = 9001
NEBO_MAX_STENCIL_POINTS = 27

# opcodes (enum nebo_opcode)
OPCODES = [
    "SCALAR", "FIELD", "MASK", "ADD", "SUB", "MUL", "DIV", "NEG", "MAX", "MIN",
    "ABS", "SQRT", "SQUARE", "CUBE", "POW", "POWINT", "EXP", "LOG", "LOG10", "SIN",
    "COS", "TAN", "TANH", "ERF", "ERFC", "EQ", "NE", "LT", "LE", "GT", "GE",
    "AND", "OR", "NOT", "STENCIL", "COND", "MASK_SHIFT", "INV_ERF", "INV_ERFC",
]
OP = {name: i for i, name in enumerate(OPCODES)}

STENCIL_WEIGHTED = 0
STENCIL_SUM = 1
STENCIL_EDGELESS = 2

RED_MAX, RED_MIN, RED_SUM, RED_NORM = 0, 1, 2, 3

KERNEL_GENERIC, KERNEL_DIVKGRAD, KERNEL_LAPL_EULER = 0, 1, 2

I3 = C.c_int32 * 3


class NeboError(RuntimeError):
    """Raised for a non-zero status of the C-ABI."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class FieldT(C.Structure):
    _fields_ = [
        ("values", C.c_void_p),
        ("glob", I3), ("offset", I3), ("extent", I3),
        ("ghost_minus", I3), ("ghost_plus", I3),
        ("valid_minus", I3), ("valid_plus", I3),
        ("has_bc", I3), ("bc_extra", I3),
        ("device", C.c_int32), ("pad_", C.c_int32),
    ]


class MaskT(C.Structure):
    _fields_ = [
        ("bits", C.c_void_p),
        ("glob", I3), ("offset", I3), ("extent", I3),
        ("ghost_minus", I3), ("ghost_plus", I3),
        ("valid_minus", I3), ("valid_plus", I3),
        ("has_bc", I3), ("bc_extra", I3),
    ]


class NodeT(C.Structure):
    _fields_ = [
        ("op", C.c_int32), ("a", C.c_int32), ("b", C.c_int32), ("c", C.c_int32),
        ("idx", C.c_int32), ("pad_", C.c_int32), ("value", C.c_double),
    ]


class StencilT(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("npts", C.c_int32),
        ("pts", (C.c_int32 * 3) * NEBO_MAX_STENCIL_POINTS),
        ("coefs", C.c_double * NEBO_MAX_STENCIL_POINTS),
    ]


class ExprT(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("nodes", C.POINTER(NodeT)),
        ("n_fields", C.c_int32), ("fields", C.POINTER(FieldT)),
        ("n_stencils", C.c_int32), ("stencils", C.POINTER(StencilT)),
        ("n_masks", C.c_int32), ("masks", C.POINTER(MaskT)),
    ]


class LimitsT(C.Structure):
    _fields_ = [("lo", I3), ("hi", I3)]


class AxisCoefsT(C.Structure):
    _fields_ = [
        ("interp", (C.c_double * 2) * 3),
        ("grad", (C.c_double * 2) * 3),
        ("div", (C.c_double * 2) * 3),
    ]


class ParticleGeomT(C.Structure):
    _fields_ = [("dx", C.c_double), ("xlo", C.c_double), ("dy", C.c_double), ("ylo", C.c_double),
                ("dz", C.c_double), ("zlo", C.c_double)]


class PatchT(C.Structure):
    _fields_ = [("a", C.c_void_p), ("b", C.c_void_p), ("n", I3), ("pad_", C.c_int32)]


P = C.POINTER
_vp = C.c_void_p
_SIGS = {
    "nebo_last_error": (C.c_char_p, []),
    "nebo_set_report_backend": (C.c_int, [C.c_int]),
    "nebo_last_backend": (C.c_char_p, []),
    "nebo_device_count": (C.c_int, [P(C.c_int)]),
    "nebo_init": (C.c_int, [C.c_int]),
    "nebo_alloc": (C.c_int, [C.c_int, C.c_size_t, P(_vp)]),
    "nebo_free": (C.c_int, [C.c_int, _vp]),
    "nebo_pool_trim": (C.c_int, [C.c_int]),
    "nebo_pool_stats": (C.c_int, [C.c_int, P(C.c_size_t), P(C.c_size_t)]),
    "nebo_memcpy": (C.c_int, [_vp, _vp, C.c_size_t, C.c_int, _vp]),
    "nebo_memcpy_peer": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_size_t, _vp]),
    "nebo_memset": (C.c_int, [_vp, C.c_int, C.c_size_t, _vp]),
    "nebo_stream_sync": (C.c_int, [_vp]),
    "nebo_device_sync": (C.c_int, []),
    "nebo_assign_limits": (C.c_int, [P(FieldT), P(ExprT), C.c_int, P(LimitsT), P(C.c_int)]),
    "nebo_assign": (C.c_int, [P(FieldT), P(ExprT), C.c_int, _vp]),
    "nebo_assign_generic": (C.c_int, [P(FieldT), P(ExprT), C.c_int, _vp]),
    "nebo_assign_plan": (C.c_int, [P(FieldT), P(ExprT), C.c_int, P(C.c_int)]),
    "nebo_lower_stats": (C.c_int, [P(ExprT), P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
    "nebo_masked_assign": (C.c_int, [P(FieldT), _vp, C.c_int32, P(ExprT), _vp]),
    "nebo_reduce": (C.c_int, [P(ExprT), C.c_int, C.c_int, P(C.c_double), _vp]),
    "nebo_reduce_async": (C.c_int, [P(ExprT), C.c_int, C.c_int, _vp, _vp]),
    "nebo_reduce_async_noalloc": (C.c_int, [P(ExprT), C.c_int, C.c_int, _vp, _vp]),
    "nebo_fill_ghosts_periodic": (C.c_int, [P(FieldT), _vp]),
    "nebo_divkgrad": (C.c_int, [P(FieldT), P(FieldT), P(FieldT), P(AxisCoefsT), C.c_int, _vp]),
    "nebo_lapl_euler": (C.c_int, [P(FieldT), P(FieldT), P(AxisCoefsT), C.c_double, _vp]),
    "nebo_stencil_tiles": (C.c_int, [P(C.c_int32)] * 5),
    "nebo_stream_triad": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int64, C.c_int32, _vp]),
    "nebo_stencil_set_tile": (C.c_int, [C.c_int, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "nebo_stencil_get_tile": (C.c_int, [C.c_int] + [P(C.c_int32)] * 4),
    "nebo_stencil_select_tile": (C.c_int, [C.c_int, C.c_int, C.c_int32, C.c_int32, C.c_int32, P(C.c_int32),
                                           P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
    "nebo_stencil_autotune": (C.c_int, [C.c_int, P(FieldT), P(FieldT), P(FieldT), P(AxisCoefsT), _vp,
                                        P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_float)]),
    "nebo_batch_create": (C.c_int, [P(PatchT), C.c_int32, P(AxisCoefsT), C.c_double, P(C.c_int32), P(_vp)]),
    "nebo_batch_destroy": (C.c_int, [_vp]),
    "nebo_batch_step": (C.c_int, [_vp, C.c_int, _vp]),
    "nebo_batch_step_per_patch": (C.c_int, [_vp, C.c_int, _vp]),
    "nebo_batch_graph_build": (C.c_int, [_vp, C.c_int32, _vp]),
    "nebo_batch_graph_launch": (C.c_int, [_vp, _vp]),
    "nebo_batch_autotune": (C.c_int, [_vp, P(C.c_int32), _vp]),
    "nebo_batch_info": (C.c_int, [_vp, P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
    "nebo_diffusion_step": (C.c_int, [P(FieldT), P(FieldT), P(FieldT), P(AxisCoefsT), C.c_double, _vp]),
    "nebo_diffusion_adjoint_step": (C.c_int, [P(FieldT), _vp, P(FieldT), P(FieldT), P(FieldT),
                                              P(AxisCoefsT), C.c_double, _vp]),
    "nebo_diffusion_forward": (C.c_int, [P(FieldT), P(FieldT), P(FieldT), P(AxisCoefsT), C.c_double,
                                         C.c_int32, C.c_int32, _vp, _vp]),
    "nebo_diffusion_backward": (C.c_int, [P(FieldT), _vp, P(FieldT), P(FieldT), P(AxisCoefsT),
                                          C.c_double, C.c_int32, C.c_int32, _vp, _vp, _vp]),
    "nebo_transport_tune": (C.c_int, [C.c_int32, C.c_int32]),
    "nebo_transport_schedule": (C.c_int, [P(FieldT), P(C.c_int32), P(C.c_int32)]),
    "nebo_transport_stage": (C.c_int, [P(FieldT), P(FieldT), P(FieldT), P(FieldT), P(FieldT), P(AxisCoefsT),
                                       C.c_double, C.c_int, _vp]),
    "nebo_transport_adjoint_stage": (C.c_int, [P(FieldT), _vp, _vp, P(FieldT), _vp, P(FieldT), P(FieldT),
                                               P(FieldT), P(AxisCoefsT), C.c_double, C.c_int, _vp]),
    "nebo_transport_forward": (C.c_int, [P(FieldT), P(FieldT), P(FieldT), P(FieldT), P(AxisCoefsT),
                                         C.c_double, C.c_int32, C.c_int32, _vp, _vp]),
    "nebo_transport_backward": (C.c_int, [P(FieldT), _vp, P(FieldT), P(FieldT), P(FieldT), P(AxisCoefsT),
                                          C.c_double, C.c_int32, C.c_int32, _vp, _vp, _vp]),
    "nebo_host_alloc": (C.c_int, [C.c_size_t, P(_vp)]),
    "nebo_set_fused_jit": (C.c_int, [C.c_int]),
    "nebo_fused_jit_stats": (C.c_int, [P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
    "nebo_fused_jit_tune": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "nebo_fused_jit_column": (C.c_int, [C.c_int32, C.c_int32, C.c_int32]),
    "nebo_fused_source": (C.c_int, [P(ExprT), C.c_int32, C.c_char_p, C.c_int32, P(C.c_int32)]),
    "nebo_fused_column_launch": (C.c_int, [P(FieldT), P(ExprT), C.c_int, C.c_char_p, C.c_int32, P(C.c_int32), _vp,
                                           C.c_int32, P(C.c_int32), P(C.c_int32)]),
    "nebo_host_free": (C.c_int, [_vp]),
    "nebo_particle_to_cell": (C.c_int, [P(FieldT), _vp, _vp, _vp, _vp, _vp, C.c_int64, P(ParticleGeomT), _vp,
                                        _vp]),
    "nebo_cell_to_particle": (C.c_int, [_vp, P(FieldT), _vp, _vp, _vp, _vp, C.c_int64, P(ParticleGeomT), _vp,
                                        _vp]),
    "nebo_point_map_create": (C.c_int, [_vp, C.c_int64, C.c_int, P(_vp)]),
    "nebo_point_map_destroy": (C.c_int, [_vp]),
    "nebo_point_to_field": (C.c_int, [_vp, P(FieldT), _vp, _vp]),
    "nebo_field_to_point": (C.c_int, [_vp, _vp, P(FieldT), _vp]),
}

_lib = None


def header_symbols() -> list[str]:
    """Every function declared in include/nebo_hip.h (parsed from the header)."""
    import re
    hdr = os.path.normpath(os.path.join(_HERE, "..", "..", "..", "include", "nebo_hip.h"))
    text = open(hdr).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(nebo_\w+)\s*\(", text, re.M)))


def lib() -> C.CDLL:
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NeboError(NEBO_ERR_UNSUPPORTED,
                            f"Nebo engine library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != NEBO_OK:
        msg = lib().nebo_last_error().decode(errors="replace")
        raise NeboError(rc, msg or f"Nebo error code {rc}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
