"""Pins the oracle's adjoint of the periodic diffusion step (built from separate
grad / interp / div passes and their transposes) by mathematics alone:
  * dot-product identity <M x, y> == <x, M^T y> for the linear map in phi
  * central finite differences of L(k) = 1/2 |step(phi; k) - t|^2 along random
    periodic directions, and of L(phi)
A transposed operand, a dropped term or a sign error fails one of these."""
import math

import numpy as np
import pytest

import nebo_hip as nh
from nebo_hip import inputs
import oracle_lib as orc


def hf(L, data):
    f = nh.Field.empty(L, "cpu")
    f.host[:] = np.asarray(data, dtype=np.float64).reshape(-1)
    return f


def interior(a, L):
    return a.reshape(L.glob[::-1])[1:-1, 1:-1, 1:-1]


def periodic(L, rng, fn):
    n = L.interior
    a = np.zeros(L.glob[::-1])
    a[1:-1, 1:-1, 1:-1] = fn(rng, (n[2], n[1], n[0]))
    return inputs.periodic_fill(a, 1).reshape(-1)


@pytest.fixture(params=[(6, 6, 6), (7, 5, 4)])
def setup(request):
    n = request.param
    L = nh.Layout.build(n, 1)
    db = nh.build_stencils(*n, 1.0, 1.0, 1.0)
    c = nh.axis_coefs(db)
    rng = np.random.default_rng(sum(n))
    k = periodic(L, rng, lambda r, s: r.uniform(0.5, 2.0, s))
    dt = 0.1 / (max(n) ** 2 * 2.0)
    return L, c, rng, k, dt


def step(L, c, phi, k, dt):
    out = hf(L, np.zeros(L.npts))
    orc.diffusion_step(out, hf(L, phi), hf(L, k), c, dt)
    return out.host


def adj(L, c, lam, phi, k, dt):
    out = hf(L, np.zeros(L.npts))
    gk = np.zeros(L.npts)
    orc.diffusion_adjoint_step(out, gk, hf(L, lam), hf(L, phi), hf(L, k), c, dt)
    return out.host, gk


def test_dot_product_identity(setup):
    L, c, rng, k, dt = setup
    x = periodic(L, rng, lambda r, s: r.standard_normal(s))
    y = periodic(L, rng, lambda r, s: r.standard_normal(s))
    Mx = step(L, c, x, k, dt)
    MTy, _ = adj(L, c, y, x, k, dt)
    lhs = math.fsum((interior(Mx, L) * interior(y, L)).ravel())
    rhs = math.fsum((interior(x, L) * interior(MTy, L)).ravel())
    assert abs(lhs - rhs) <= 1e-13 * (abs(lhs) + abs(rhs))


def test_gradient_wrt_k_matches_central_differences(setup):
    L, c, rng, k, dt = setup
    phi = periodic(L, rng, lambda r, s: r.standard_normal(s))
    target = periodic(L, rng, lambda r, s: r.standard_normal(s))

    def loss(kk):
        d = interior(step(L, c, phi, kk, dt), L) - interior(target, L)
        return 0.5 * math.fsum((d * d).ravel())

    lam = step(L, c, phi, k, dt) - target
    lam = inputs.periodic_fill(lam.reshape(L.glob[::-1]).copy(), 1).reshape(-1)
    _, gk = adj(L, c, lam, phi, k, dt)
    eps = 1e-6
    for _ in range(6):
        v = periodic(L, rng, lambda r, s: r.standard_normal(s))
        fd = (loss(k + eps * v) - loss(k - eps * v)) / (2 * eps)
        an = math.fsum((interior(gk, L) * interior(v, L)).ravel())
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)


def test_gradient_wrt_phi_matches_central_differences(setup):
    L, c, rng, k, dt = setup
    phi = periodic(L, rng, lambda r, s: r.standard_normal(s))
    target = periodic(L, rng, lambda r, s: r.standard_normal(s))

    def loss(p):
        d = interior(step(L, c, p, k, dt), L) - interior(target, L)
        return 0.5 * math.fsum((d * d).ravel())

    lam = step(L, c, phi, k, dt) - target
    lam = inputs.periodic_fill(lam.reshape(L.glob[::-1]).copy(), 1).reshape(-1)
    g, _ = adj(L, c, lam, phi, k, dt)
    eps = 1e-6
    for _ in range(4):
        v = periodic(L, rng, lambda r, s: r.standard_normal(s))
        fd = (loss(phi + eps * v) - loss(phi - eps * v)) / (2 * eps)
        an = math.fsum((interior(g, L) * interior(v, L)).ravel())
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)


def test_ghosts_of_adjoint_are_periodic(setup):
    L, c, rng, k, dt = setup
    y = periodic(L, rng, lambda r, s: r.standard_normal(s))
    x = periodic(L, rng, lambda r, s: r.standard_normal(s))
    out, gk = adj(L, c, y, x, k, dt)
    ref = inputs.periodic_fill(out.reshape(L.glob[::-1]).copy(), 1)
    assert np.array_equal(out.reshape(L.glob[::-1]), ref)
    # gk is written on interior cells only
    g = gk.reshape(L.glob[::-1])
    assert np.all(g[0] == 0) and np.all(g[-1] == 0) and np.all(g[:, 0] == 0) and np.all(g[:, :, -1] == 0)
