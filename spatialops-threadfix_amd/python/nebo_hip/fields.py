"""Field layouts, fields and masks for the Python binding.

A :class:`Layout` holds exactly the metadata of a reference SpatialField:
MemoryWindow (glob, offset, extent; spatialops/structured/MemoryWindow.h:63-321),
GhostData (total and valid ghost layers; GhostData.h:54-257) and BoundaryCellInfo
(has_bc / has_extra; BoundaryCellInfo.h:45-158).  A :class:`Field` keeps a host
copy (numpy) and/or a GPU copy (torch tensor) of the whole glob block, like the
reference's per-device map in FieldInfo.h:91-416.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _abi

# Location traits: offset and BCExtra of spatialops/structured/FVStaggeredLocationTypes.h:63-126
LOCATIONS = {
    "SVol": ((0, 0, 0), (0, 0, 0)),
    "SSurfX": ((-1, 0, 0), (1, 0, 0)), "SSurfY": ((0, -1, 0), (0, 1, 0)), "SSurfZ": ((0, 0, -1), (0, 0, 1)),
    "XVol": ((-1, 0, 0), (1, 0, 0)),
    "XSurfX": ((0, 0, 0), (0, 0, 0)), "XSurfY": ((-1, -1, 0), (0, 1, 0)), "XSurfZ": ((-1, 0, -1), (0, 0, 1)),
    "YVol": ((0, -1, 0), (0, 1, 0)),
    "YSurfX": ((-1, -1, 0), (1, 0, 0)), "YSurfY": ((0, 0, 0), (0, 0, 0)), "YSurfZ": ((0, -1, -1), (0, 0, 1)),
    "ZVol": ((0, 0, -1), (0, 0, 1)),
    "ZSurfX": ((-1, 0, -1), (1, 0, 0)), "ZSurfY": ((0, -1, -1), (0, 1, 0)), "ZSurfZ": ((0, 0, 0), (0, 0, 0)),
    "SingleValue": ((0, 0, 0), (0, 0, 0)),
}

# FaceTypes<Vol> (FVStaggeredFieldTypes.h)
FACE_TYPES = {
    "SVol": ("SSurfX", "SSurfY", "SSurfZ"),
    "XVol": ("XSurfX", "XSurfY", "XSurfZ"),
    "YVol": ("YSurfX", "YSurfY", "YSurfZ"),
    "ZVol": ("ZSurfX", "ZSurfY", "ZSurfZ"),
}


def _v3(x):
    if isinstance(x, int):
        return (x, x, x)
    return tuple(int(v) for v in x)


@dataclass(frozen=True)
class Layout:
    """Window, ghosts and boundary info of one field."""
    glob: tuple
    offset: tuple
    extent: tuple
    ghost_minus: tuple
    ghost_plus: tuple
    valid_minus: tuple
    valid_plus: tuple
    has_bc: tuple
    bc_extra: tuple
    loc: str = "SVol"

    @staticmethod
    def build(interior, ghost=1, loc="SVol", bc=(False, False, False)):
        """get_window_with_ghost(interior, GhostData(ghost), BoundaryCellInfo::build<Loc>(bc)),
        with the extent-limited ghosts of the SpatialField constructor (SpatialField.h:155-165)."""
        n = _v3(interior)
        if isinstance(ghost, int):
            gm, gp = (ghost,) * 3, (ghost,) * 3
        else:
            gm, gp = _v3(ghost[0]), _v3(ghost[1])
        bc = (bool(bc),) * 3 if isinstance(bc, (bool, int)) else tuple(bool(b) for b in bc)
        extra_loc = LOCATIONS[loc][1]
        extra = tuple(extra_loc[d] if bc[d] else 0 for d in range(3))
        glob = tuple(n[d] + gm[d] + gp[d] + extra[d] if n[d] > 1 else 1 for d in range(3))
        # limit_by_extent: axes of extent 1 carry no ghosts and no extra cells
        gm = tuple(gm[d] if glob[d] > 1 else 0 for d in range(3))
        gp = tuple(gp[d] if glob[d] > 1 else 0 for d in range(3))
        extra = tuple(extra[d] if glob[d] > 1 else 0 for d in range(3))
        return Layout(glob=glob, offset=(0, 0, 0), extent=glob, ghost_minus=gm, ghost_plus=gp,
                      valid_minus=gm, valid_plus=gp, has_bc=tuple(int(b) for b in bc),
                      bc_extra=extra, loc=loc)

    def like(self, loc):
        """A field of another location on the same window (face fields share the
        volume window when there is no BC extra cell)."""
        return Layout.build(self.interior, (self.ghost_minus, self.ghost_plus), loc,
                            tuple(bool(b) for b in self.has_bc))

    def create_new(self, loc, use_valid=False):
        """Layout of a field of location ``loc`` next to a field of this layout
        (create_new_memory_window / create_new_boundary_cell_info,
        SpatialField.h:720-755): the window grows by the difference of the two
        locations' BC extra cells on the axes that carry a boundary.  Ghosts are
        this layout's total ghosts, or its valid ghosts when ``use_valid``."""
        new_extra, old_extra = LOCATIONS[loc][1], LOCATIONS[self.loc][1]
        bc = tuple(int(b) for b in self.has_bc)
        inc = tuple(bc[d] * (new_extra[d] - old_extra[d]) for d in range(3))
        gm = self.valid_minus if use_valid else self.ghost_minus
        gp = self.valid_plus if use_valid else self.ghost_plus
        return Layout(glob=tuple(self.glob[d] + inc[d] for d in range(3)), offset=self.offset,
                      extent=tuple(self.extent[d] + inc[d] for d in range(3)), ghost_minus=gm,
                      ghost_plus=gp, valid_minus=gm, valid_plus=gp, has_bc=bc,
                      bc_extra=tuple(new_extra[d] if bc[d] else 0 for d in range(3)), loc=loc)

    @property
    def interior(self):
        return tuple(self.extent[d] - self.ghost_minus[d] - self.ghost_plus[d] - self.bc_extra[d]
                     for d in range(3))

    @property
    def npts(self):
        return self.glob[0] * self.glob[1] * self.glob[2]

    def with_valid(self, minus, plus):
        from dataclasses import replace
        return replace(self, valid_minus=_v3(minus), valid_plus=_v3(plus))

    def desc(self, values_ptr: int, device: int) -> _abi.FieldT:
        f = _abi.FieldT()
        f.values = values_ptr
        for name in ("glob", "offset", "extent", "ghost_minus", "ghost_plus",
                     "valid_minus", "valid_plus", "has_bc", "bc_extra"):
            getattr(f, name)[:] = list(getattr(self, name))
        f.device = device
        return f


class Field:
    """One field: a host numpy copy and/or a GPU torch copy of the glob block."""

    def __init__(self, layout: Layout, host=None, gpu=None, device: int = 0, name: str = ""):
        self.layout = layout
        self.host = host
        self.gpu = gpu
        self.device = device
        self.name = name

    @staticmethod
    def like(prototype: "Field", loc=None, where=None, device=None, name=""):
        """SpatialFieldStore::get<loc>(prototype) (SpatialFieldStore.h:240-257): an
        uninitialised field of location ``loc`` (default: the prototype's) shaped
        after ``prototype``, where the prototype lives.  GPU storage comes from
        torch's caching allocator, which keeps freed blocks for reuse as the
        reference's store does."""
        L = prototype.layout.create_new(loc or prototype.layout.loc)
        if where is None:
            where = "gpu" if prototype.gpu is not None else "cpu"
        return Field.empty(L, where, prototype.device if device is None else device, name)

    @staticmethod
    def empty(layout: Layout, where="gpu", device=0, name=""):
        n = layout.npts
        if where == "gpu":
            return Field(layout, gpu=torch.empty(n, dtype=torch.float64, device=f"cuda:{device}"),
                         device=device, name=name)
        return Field(layout, host=np.empty(n, dtype=np.float64), device=device, name=name)

    @staticmethod
    def from_numpy(layout: Layout, arr, where="both", device=0, name=""):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        assert a.size == layout.npts, (a.size, layout.npts)
        f = Field(layout, host=a.copy(), device=device, name=name)
        if where in ("both", "gpu"):
            f.gpu = torch.from_numpy(f.host).to(f"cuda:{device}")
        if where == "gpu":
            f.host = None
        return f

    def to_gpu(self):
        self.gpu = torch.from_numpy(self.host).to(f"cuda:{self.device}")
        return self

    def to_host(self):
        self.host = self.gpu.detach().cpu().numpy().copy()
        return self

    def desc(self, where="gpu") -> _abi.FieldT:
        if where == "gpu":
            return self.layout.desc(self.gpu.data_ptr(), self.device)
        return self.layout.desc(self.host.ctypes.data, _abi.NEBO_CPU_INDEX)

    def array(self, where="host"):
        """The glob block as a (z, y, x) array."""
        g = self.layout.glob
        src = self.host if where == "host" else self.gpu.detach().cpu().numpy()
        return src.reshape(g[2], g[1], g[0])


class Mask:
    """SpatialMask(prototype, points): a bit field over the prototype's glob block
    (spatialops/structured/SpatialMask.h:99-181, BitField.h:291-362).  Points are
    interior-relative (x, y, z) triples; bit b of word b//32 marks flat index b."""

    def __init__(self, layout: Layout, points, device=0):
        self.layout = layout
        self.points = np.asarray(points, dtype=np.int32).reshape(-1, 3)
        g = layout.glob
        npts = g[0] * g[1] * g[2]
        words = np.zeros((npts + 31) // 32, dtype=np.uint32)
        p = self.points
        gx = layout.offset[0] + layout.ghost_minus[0] + p[:, 0]
        gy = layout.offset[1] + layout.ghost_minus[1] + p[:, 1]
        gz = layout.offset[2] + layout.ghost_minus[2] + p[:, 2]
        pos = gx.astype(np.int64) + g[0] * (gy.astype(np.int64) + g[1] * gz.astype(np.int64))
        np.bitwise_or.at(words, pos // 32, (np.uint32(1) << (pos % 32).astype(np.uint32)))
        self.words = words
        self.device = device
        self.gpu = None

    @staticmethod
    def build(prototype, loc, points, device=0):
        """SpatialMask<loc>::build(prototype, points) (SpatialMask.h:173-181): a mask of
        location ``loc`` on the window a field of that location would have next to
        ``prototype`` (a Field or Layout), with the prototype's valid ghosts."""
        L = prototype.layout if isinstance(prototype, Field) else prototype
        return Mask(L.create_new(loc, use_valid=True), points, device)

    def add_consumer(self, device=0):
        """Copy the bit field to a GPU (BitField::add_consumer, BitField.h:466-481)."""
        self.device = device
        self.gpu = torch.from_numpy(self.words.view(np.int32)).to(f"cuda:{device}")
        return self

    def desc(self, where="gpu") -> _abi.MaskT:
        m = _abi.MaskT()
        m.bits = self.gpu.data_ptr() if where == "gpu" else self.words.ctypes.data
        L = self.layout
        for name in ("glob", "offset", "extent", "ghost_minus", "ghost_plus",
                     "valid_minus", "valid_plus", "has_bc", "bc_extra"):
            getattr(m, name)[:] = list(getattr(L, name))
        return m
