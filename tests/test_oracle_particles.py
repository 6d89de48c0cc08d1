"""Particle <-> cell interpolation and point fields on the CPU (oracle).

Pinning:
  * the known answers of spatialops/particles/test/TestParticleOperators.cpp
    (c2p = 32.5, p2c = 1,4,4,4,4,3,0,0,0,0) and of
    spatialops/pointfield/test/testPointField.cpp, exactly;
  * an independent vectorised restatement of the overlap geometry (every
    particle against every cell, overlaps from interval intersection) agrees to
    rounding on random 1-, 2- and 3-D particle sets;
  * closed forms: P2C conserves the particle total, C2P of a constant field is
    that constant, a particle that is exactly one cell moves its value to or
    from that cell without rounding.
"""
import numpy as np
import pytest

import nebo_hip as nh
import oracle_lib as orc
from particle_cases import random_particles, reference_case


def host(L, a=None):
    f = nh.Field.empty(L, "cpu")
    f.host[:] = 0.0 if a is None else np.asarray(a).reshape(-1)
    return f


def test_reference_particle_case():
    L, x, f, geom = reference_case()
    cells = host(L, f)
    ptmp = orc.cell_to_particle(cells, np.array([3.25]), None, None, np.array([5.0]), geom, 1)
    assert ptmp[0] == 32.5
    ctmp = host(L)
    orc.particle_to_cell(ctmp, np.array([20.0]), np.array([3.25]), None, None, np.array([5.0]), geom)
    assert list(ctmp.host[3:13]) == [1, 4, 4, 4, 4, 3, 0, 0, 0, 0]


def test_reference_point_field_case():
    f = np.array([0.1 + i for i in range(10)])
    f[2], f[4] = 1.234, 3.456
    ix = [4, 2]
    f2 = orc.field_to_point(f, ix)
    assert list(f2) == [3.456, 1.234]
    orc.point_to_field(f, np.array([-1.234, -3.456]), ix)
    assert f[2] == -3.456 and f[4] == -1.234


def brute_overlaps(L, coords, size, geom):
    """Fraction of each particle's cube in each window cell, every pair at once."""
    d = [geom[0], geom[2], geom[4]]
    o = [geom[1], geom[3], geom[5]]
    frac = np.ones((len(size),) + tuple(L.extent[::-1]))
    for a in range(3):
        h = d[a] if d[a] > 0 else 1.0
        lo = o[a] if d[a] > 0 else 0.0
        n = L.extent[a]
        face = lo - h * L.ghost_minus[a] - h / 2 + h * np.arange(n)
        if coords[a] is None:
            t = (0.0 - (lo - h * L.ghost_minus[a] - h / 2)) / h
            w = np.zeros((len(size), n))
            w[:, int(t)] = 1.0
        else:
            plo = coords[a] - size / 2
            phi = plo + size
            w = np.clip(np.minimum(face[None, :] + h, phi[:, None]) - np.maximum(face[None, :], plo[:, None]),
                        0.0, None) / size[:, None]
        shape = [len(size), 1, 1, 1]
        shape[3 - a] = n
        frac = frac * w.reshape(shape)
    return frac   # (p, z, y, x) over the window


@pytest.mark.parametrize("axes", [(0,), (0, 1), (0, 1, 2), (0, 2)])
def test_matches_vectorised_geometry(axes):
    n = tuple(12 if a in axes else 1 for a in range(3))
    L = nh.Layout.build(n, 2, "SVol")
    src, coords, size, geom = random_particles(L, 60, 5, axes=axes, dx=(0.5, 0.25, 2.0), lo=(0.25, 0.125, 1.0))
    frac = brute_overlaps(L, coords, size, geom)
    cells = host(L)
    orc.particle_to_cell(cells, src, *coords, size, geom)
    want = np.einsum("p,pzyx->zyx", src, frac)
    got = cells.array()
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(src).sum())
    f = host(L, np.random.default_rng(1).standard_normal(L.npts))
    c2p = orc.cell_to_particle(f, *coords, size, geom, len(size))
    want = np.einsum("pzyx,zyx->p", frac, f.array())
    assert np.allclose(c2p, want, rtol=1e-12, atol=1e-12)


def test_closed_forms():
    L = nh.Layout.build((10, 9, 8), 1, "SVol")
    src, coords, size, geom = random_particles(L, 500, 9, dx=(1.0, 1.0, 1.0), lo=(0.5, 0.5, 0.5))
    cells = host(L)
    orc.particle_to_cell(cells, src, *coords, size, geom)
    assert abs(cells.host.sum() - src.sum()) <= 1e-12 * np.abs(src).sum()
    const = host(L, np.full(L.npts, 3.0))
    c2p = orc.cell_to_particle(const, *coords, size, geom, len(size))
    assert np.allclose(c2p, 3.0, rtol=1e-13, atol=0)   # the reference asserts |1 - sum| < 1e-10
    # a unit particle sitting exactly on cell (4, 3, 2): one cell, fraction 1
    one = [np.array([4.5]), np.array([3.5]), np.array([2.5])]
    vals = host(L, np.random.default_rng(2).standard_normal(L.npts))
    got = orc.cell_to_particle(vals, *one, np.array([1.0]), geom, 1)
    assert got[0] == vals.array()[2 + 1, 3 + 1, 4 + 1]
    cells = host(L)
    orc.particle_to_cell(cells, np.array([0.7]), *one, np.array([1.0]), geom)
    assert cells.array()[3, 4, 5] == 0.7 and np.count_nonzero(cells.host) == 1


def test_particle_below_the_window_touches_nothing():
    L = nh.Layout.build((10, 1, 1), 1, "SVol")
    geom = (1.0, 0.5, -1.0, 0.0, -1.0, 0.0)
    cells = host(L)
    orc.particle_to_cell(cells, np.array([5.0]), np.array([-3.0]), None, None, np.array([1.0]), geom)
    assert not cells.host.any()
    vals = host(L, np.arange(L.npts, dtype=float))
    assert orc.cell_to_particle(vals, np.array([-3.0]), None, None, np.array([1.0]), geom, 1)[0] == 0.0
