"""Inputs shared by the CPU and GPU particle tests: the case of
spatialops/particles/test/TestParticleOperators.cpp and random particle sets."""
from __future__ import annotations

import numpy as np

import nebo_hip as nh


def reference_case():
    """dim (10,1,1), length (10,1,1), 3 ghost cells, SVol; xcoord from
    Grid::set_coord<XDIR>, f = 10*xcoord, one particle at 3.25 of size 5 carrying
    20 (TestParticleOperators.cpp:27-80).  Expected: c2p gives 32.5 and p2c puts
    1, 4, 4, 4, 4, 3, 0, 0, 0, 0 on the interior cells."""
    L = nh.Layout.build((10, 1, 1), 3, "SVol")
    dx = 10.0 / 10
    # Grid::set_coord: dx*(i - ghost) + 0.5*dx*(1 + stagger) over the window (Grid.cpp:58-102)
    x = np.array([dx * (i - 3) + 0.5 * dx * (1 + 0) for i in range(L.glob[0])])
    f = 10 * x
    geom = (x[1] - x[0], x[3], -1.0, 0.0, -1.0, 0.0)
    return L, x, f, geom


def random_particles(L, n, seed, axes=(0, 1, 2), max_size=2.5, dx=(1.0, 1.0, 1.0), lo=(0.5, 0.5, 0.5)):
    """n particles inside the interior of layout L (cell spacing dx, first interior
    centre lo), cube edges up to max_size cells."""
    rng = np.random.default_rng(seed)
    interior = L.interior
    size = rng.uniform(0.2, max_size, n) * min(dx[a] for a in axes)
    coords = []
    for a in range(3):
        if a not in axes:
            coords.append(None)
            continue
        x0 = lo[a] - 0.5 * dx[a]
        x1 = x0 + interior[a] * dx[a]
        coords.append(rng.uniform(x0 + size / 2, x1 - size / 2))
    src = rng.standard_normal(n)
    geom = (dx[0] if 0 in axes else -1.0, lo[0], dx[1] if 1 in axes else -1.0, lo[1],
            dx[2] if 2 in axes else -1.0, lo[2])
    return src, coords, size, geom
