"""Field layouts of temporaries and masks built next to a prototype
(create_new_memory_window, SpatialField.h:720-755), checked the way
spatialops/structured/test/testField.cpp:246-296 checks SpatialFieldStore: a
window derived from a field of another location equals the window built
directly for the new location.  CPU only."""
import pytest

import nebo_hip as nh

PAIRS = [("SVol", "SVol", (0, 0, 0)), ("SVol", "SVol", (1, 1, 1)), ("SVol", "SSurfX", (0, 0, 0)),
         ("SSurfX", "SVol", (0, 0, 0)), ("SVol", "SSurfY", (1, 1, 1)), ("SVol", "SSurfY", (0, 0, 0)),
         ("SVol", "SSurfX", (1, 1, 1)), ("SSurfX", "SVol", (1, 1, 1)), ("SVol", "SSurfZ", (1, 1, 1)),
         ("SVol", "SSurfZ", (0, 0, 0)), ("XSurfX", "SVol", (0, 0, 0)), ("XSurfX", "ZVol", (0, 0, 0)),
         ("YSurfZ", "ZSurfX", (0, 0, 0)), ("XVol", "YSurfX", (1, 0, 1)), ("ZVol", "XSurfZ", (1, 1, 1))]


@pytest.mark.parametrize("loc1,loc2,bc", PAIRS)
def test_store_windows(loc1, loc2, bc):
    dim = (30, 40, 50)
    L1 = nh.Layout.build(dim, 1, loc1, bc)
    L2 = nh.Layout.build(dim, 1, loc2, bc)
    for derived in (L1.create_new(loc2), L2.create_new(loc2)):
        assert (derived.glob, derived.offset, derived.extent) == (L2.glob, L2.offset, L2.extent)
        assert derived.bc_extra == L2.bc_extra and derived.loc == loc2


def test_mask_build_uses_valid_ghosts():
    L = nh.Layout.build((6, 5, 4), 2, "SVol", (True, True, True)).with_valid(1, 1)
    m = nh.Mask.build(L, "SSurfX", [(0, 0, 0)])
    assert m.layout.glob == (11, 9, 8) and m.layout.ghost_minus == (1, 1, 1)
    assert m.layout.valid_plus == (1, 1, 1) and m.layout.bc_extra == (1, 0, 0)


def test_field_like_on_host():
    L = nh.Layout.build((6, 5, 4), 1, "XVol", (True, False, False))
    f = nh.Field.empty(L, "cpu")
    g = nh.Field.like(f, "SVol")
    assert g.host is not None and g.layout.glob == nh.Layout.build((6, 5, 4), 1, "SVol", (True, False, False)).glob
