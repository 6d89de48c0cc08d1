"""Particle <-> cell interpolation and point fields on the GPU through the C-ABI,
against the oracle (tests/test_oracle_particles.py pins it) and the known
answers of the reference's particle and point-field tests.  Bar: bit-exact."""
import numpy as np
import pytest
import torch

import nebo_hip as nh
from nebo_hip import particles as P
import oracle_lib as orc
from particle_cases import random_particles, reference_case

pytestmark = pytest.mark.gpu


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


def both(L, a):
    return nh.Field.from_numpy(L, a, where="both")


def fetch(f):
    torch.cuda.synchronize()
    return f.gpu.cpu().numpy()


def test_reference_particle_case_on_gpu():
    L, x, f, geom = reference_case()
    c2p = P.CellToParticle(*geom)
    p2c = P.ParticleToCell(*geom)
    pc, ps = dev([3.25]), dev([5.0])
    for op in (c2p, p2c):
        op.set_coordinate_information(pc, None, None, ps)
    ptmp = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    c2p.apply_to_field(both(L, f), ptmp)
    ctmp = both(L, np.full(L.npts, 7.0))
    p2c.apply_to_field(dev([20.0]), ctmp)
    assert ptmp.item() == 32.5
    assert list(fetch(ctmp)[3:13]) == [1, 4, 4, 4, 4, 3, 0, 0, 0, 0]


@pytest.mark.parametrize("axes", [(0,), (0, 1), (0, 1, 2), (1, 2)])
@pytest.mark.parametrize("npart", [1, 777, 40000])
def test_random_particles_bit_exact(axes, npart):
    n = tuple(40 if a in axes else 1 for a in range(3))
    L = nh.Layout.build(n, 2, "SVol")
    src, coords, size, geom = random_particles(L, npart, npart + len(axes), axes=axes,
                                               dx=(0.5, 0.25, 2.0), lo=(0.25, 0.125, 1.0))
    # P2C: many particles per cell, sums in particle order
    cells = both(L, np.full(L.npts, -1.0))
    skipped = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    op = P.ParticleToCell(*geom)
    op.set_coordinate_information(*[dev(c) for c in coords], dev(size))
    op.apply_to_field(dev(src), cells, skipped=skipped)
    orc.particle_to_cell(cells, src, *coords, size, geom)
    assert np.array_equal(fetch(cells), cells.host)
    assert skipped.item() == 0
    # C2P
    vals = both(L, np.random.default_rng(npart).standard_normal(L.npts))
    out = torch.zeros(npart, dtype=torch.float64, device="cuda:0")
    c2p = P.CellToParticle(*geom)
    c2p.set_coordinate_information(*[dev(c) for c in coords], dev(size))
    c2p.apply_to_field(vals, out)
    want = orc.cell_to_particle(vals, *coords, size, geom, npart)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), want)


def test_particles_outside_and_on_ghost_cells():
    """Particles that start below the window touch nothing (and are counted);
    particles over ghost cells and past the top are clipped as the reference
    clips them."""
    L = nh.Layout.build((16, 12, 1), 2, "SVol", (True, False, False))
    rng = np.random.default_rng(3)
    npart = 3000
    size = rng.uniform(0.3, 4.0, npart)
    px = rng.uniform(-4.0, 20.0, npart)
    py = rng.uniform(-4.0, 16.0, npart)
    src = rng.standard_normal(npart)
    geom = (1.0, 0.5, 1.0, 0.5, -1.0, 0.0)
    cells = both(L, np.zeros(L.npts))
    skipped = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    op = P.ParticleToCell(*geom)
    op.set_coordinate_information(dev(px), dev(py), None, dev(size))
    op.apply_to_field(dev(src), cells, skipped=skipped)
    orc.particle_to_cell(cells, src, px, py, None, size, geom)
    assert np.array_equal(fetch(cells), cells.host)
    face = 0.5 - 2.0 - 0.5
    lo_x, lo_y = (px - size / 2 - face), (py - size / 2 - face)
    assert skipped.item() == int(np.count_nonzero((lo_x <= -1.0) | (lo_y <= -1.0)))
    out = torch.zeros(npart, dtype=torch.float64, device="cuda:0")
    c2p = P.CellToParticle(*geom)
    c2p.set_coordinate_information(dev(px), dev(py), None, dev(size))
    c2p.apply_to_field(cells, out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), orc.cell_to_particle(cells, px, py, None, size, geom, npart))


def test_point_fields_on_gpu():
    L = nh.Layout.build((10, 1, 1), 0, "SVol")
    f = np.array([0.1 + i for i in range(10)])
    f[2], f[4] = 1.234, 3.456
    fld = both(L, f)
    ftp, ptf = P.FieldToPoint([4, 2]), P.PointToField([4, 2])
    f2 = torch.zeros(2, dtype=torch.float64, device="cuda:0")
    ftp.apply_to_field(fld, f2)
    assert f2.cpu().tolist() == [3.456, 1.234]
    ptf.apply_to_field(dev([-1.234, -3.456]), fld)
    got = fetch(fld)
    assert got[2] == -3.456 and got[4] == -1.234
    # repeated indices: the last point wins, as in the reference's loop
    big = nh.Layout.build((30, 20, 10), 1, "SVol")
    rng = np.random.default_rng(4)
    idx = rng.integers(0, big.npts, 5000)
    vals = rng.standard_normal(5000)
    g = both(big, rng.standard_normal(big.npts))
    host = g.host.copy()
    P.PointToField(idx).apply_to_field(dev(vals), g)
    orc.point_to_field(host, vals, idx)
    assert np.array_equal(fetch(g), host)
    out = torch.zeros(5000, dtype=torch.float64, device="cuda:0")
    P.FieldToPoint(idx).apply_to_field(g, out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), orc.field_to_point(host, idx))
    # from a locations field: nonzero entries in iteration order
    loc = np.zeros(big.npts)
    loc[[7, 300, 2999]] = 1.0
    m = P.FieldToPoint(nh.Field.from_numpy(big, loc, where="cpu"))
    out3 = torch.zeros(3, dtype=torch.float64, device="cuda:0")
    m.apply_to_field(g, out3)
    torch.cuda.synchronize()
    assert np.array_equal(out3.cpu().numpy(), host[[7, 300, 2999]])


def test_point_index_beyond_field_is_refused():
    L = nh.Layout.build((4, 4, 4), 0, "SVol")
    g = both(L, np.zeros(L.npts))
    with pytest.raises(nh.NeboError, match="beyond"):
        P.PointToField([0, L.npts]).apply_to_field(dev([1.0, 2.0]), g)
