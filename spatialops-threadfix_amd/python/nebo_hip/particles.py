"""Particle <-> cell interpolation and point fields (spatialops/particles/
ParticleOperators.h:35-395, spatialops/pointfield/PointOperators.h:28-160).

Particle fields are 1-D arrays (the reference's ParticleField has an (np, 1, 1)
window and no ghosts): torch float64 tensors on the cell field's GPU.  The
operators keep the reference's construction and ``apply_to_field`` interface.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _abi
from .api import _stream
from .fields import Field


def _ptr(t):
    if t is None:
        return None
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise TypeError("particle arrays are contiguous float64 tensors on a GPU")
    return C.c_void_p(t.data_ptr())


class _ParticleOp:
    def __init__(self, dx, xlo, dy=-1.0, ylo=0.0, dz=-1.0, zlo=0.0):
        self.geom = _abi.ParticleGeomT(float(dx), float(xlo), float(dy), float(ylo), float(dz), float(zlo))
        self.px = self.py = self.pz = self.psize = None

    def set_coordinate_information(self, px, py, pz, psize):
        """Particle coordinates (None for an axis the particles do not resolve) and
        particle sizes (the edge of each particle's cube)."""
        self.px, self.py, self.pz, self.psize = px, py, pz, psize

    def _args(self, n):
        if self.psize is None:
            raise ValueError("set_coordinate_information first")
        for t in (self.px, self.py, self.pz, self.psize):
            if t is not None and t.numel() < n:
                raise ValueError("a particle array is shorter than the particle field")
        return [_ptr(self.px), _ptr(self.py), _ptr(self.pz), _ptr(self.psize)]


class ParticleToCell(_ParticleOp):
    """ParticleToCell<CellField>(dx, xlo, dy, ylo, dz, zlo): spreads an extensive
    particle quantity over the cells each particle overlaps."""

    def apply_to_field(self, src: torch.Tensor, dest: Field, stream=None, skipped=None):
        n = src.numel()
        _abi.call("nebo_particle_to_cell", C.byref(dest.desc("gpu")), _ptr(src), *self._args(n), n,
                  C.byref(self.geom), _ptr_i32(skipped), _stream(stream))
        return dest


class CellToParticle(_ParticleOp):
    """CellToParticle<CellField>(dx, xlo, ...): the volume-weighted average of a cell
    field over each particle's cube."""

    def apply_to_field(self, src: Field, dest: torch.Tensor, stream=None, skipped=None):
        n = dest.numel()
        _abi.call("nebo_cell_to_particle", _ptr(dest), C.byref(src.desc("gpu")), *self._args(n), n,
                  C.byref(self.geom), _ptr_i32(skipped), _stream(stream))
        return dest


def _ptr_i32(t):
    if t is None:
        return None
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.int32):
        raise TypeError("the skipped counter is an int32 tensor on a GPU")
    return C.c_void_p(t.data_ptr())


class _PointMap:
    def __init__(self, indices, device=0):
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.int64).reshape(-1))
        self.n = len(idx)
        h = C.c_void_p()
        _abi.call("nebo_point_map_create", C.c_void_p(idx.ctypes.data), self.n, int(device), C.byref(h))
        self._h = h

    @staticmethod
    def indices_of(locations: Field):
        """The constructors from a field (PointOperators.h:95-104,125-134): the
        iteration positions (x fastest over the window with ghosts) of its nonzero
        entries.  The operators then use them as flat indices of the block, which
        is the same thing for a field whose window is its whole block."""
        L = locations.layout
        a = locations.array("host" if locations.host is not None else "gpu")
        o, e = L.offset, L.extent
        win = a[o[2]:o[2] + e[2], o[1]:o[1] + e[1], o[0]:o[0] + e[0]]
        return np.flatnonzero(win.reshape(-1) != 0.0).astype(np.int64)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _abi.lib().nebo_point_map_destroy(h)
            except Exception:
                pass


class PointToField(_PointMap):
    """PointToField<FieldT>(indices or a locations field): dest[index[i]] = src[i]."""

    def __init__(self, locations, device=0):
        super().__init__(self.indices_of(locations) if isinstance(locations, Field) else locations, device)

    def apply_to_field(self, src: torch.Tensor, dest: Field, stream=None):
        if src.numel() < self.n:
            raise ValueError("point field shorter than the index list")
        _abi.call("nebo_point_to_field", self._h, C.byref(dest.desc("gpu")), _ptr(src), _stream(stream))
        return dest


class FieldToPoint(_PointMap):
    """FieldToPoint<FieldT>(indices or a locations field): dest[i] = src[index[i]]."""

    def __init__(self, locations, device=0):
        super().__init__(self.indices_of(locations) if isinstance(locations, Field) else locations, device)

    def apply_to_field(self, src: Field, dest: torch.Tensor, stream=None):
        if dest.numel() < self.n:
            raise ValueError("point field shorter than the index list")
        _abi.call("nebo_field_to_point", self._h, _ptr(dest), C.byref(src.desc("gpu")), _stream(stream))
        return dest
