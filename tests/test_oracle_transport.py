"""Pins the oracle's scalar-transport functions (upwinded convection by the sign of
a face velocity plus var-coefficient diffusion, SSP-RK3, periodic):
  * the closed-form stage equals the tree walk of the fused stage expression
    (cond / stencil / pointwise nodes of the restated reference semantics), bit
    for bit, for the three stages
  * the adjoint stage satisfies the dot-product identity with the forward stage
    (linear in (phi_n, s)) and its k part matches the exact change of the
    forward stage under a change of k (the stage is affine in k)
  * the reverse sweep over a few RK3 steps matches central differences of
    L = 1/2 |phi_T - t|^2 in k and in phi_0
A dropped convective term, a wrong upwind side, a transposed stencil or a wrong
stage weight fails one of these."""
import numpy as np
import pytest

import nebo_hip as nh
from nebo_hip import inputs
from nebo_hip import workloads as W
import oracle_lib as orc


def hf(L, a):
    return nh.Field(L, host=np.asarray(a, dtype=np.float64).copy())


def interior(L, a):
    return np.asarray(a).reshape(L.glob[::-1])[1:-1, 1:-1, 1:-1]


def periodic_random(L, rng, fn):
    n = L.interior
    a = np.zeros(L.glob[::-1])
    a[1:-1, 1:-1, 1:-1] = fn(rng, (n[2], n[1], n[0]))
    return inputs.periodic_fill(a, 1).reshape(-1)


@pytest.fixture(params=[6, 9])
def problem(request):
    n = request.param
    L, Ls, ops, k, phi, u, dt = W.transport_problem(n)
    phi = inputs.periodic_fill(phi.reshape(L.glob[::-1]).copy(), 1).reshape(-1)
    us = [nh.Field(l, host=x) for l, x in zip(Ls, u)]
    c = nh.axis_coefs(nh.build_stencils(n, n, n, 1.0, 1.0, 1.0))
    return n, L, Ls, ops, k, phi, u, us, c, dt


@pytest.mark.parametrize("stage", [1, 2, 3])
def test_stage_closed_form_equals_tree_walk(problem, stage):
    n, L, Ls, ops, k, phi, u, us, c, dt = problem
    rng = np.random.default_rng(stage)
    s = periodic_random(L, rng, lambda r, sh: r.standard_normal(sh))
    lo, hi = W.upwind_ops()
    K, PN, S = hf(L, k), hf(L, phi), hf(L, s)
    s1, s2, s3 = W.rk3_stage_exprs(ops, PN, K, us, dt, lo, hi)
    expr = s1() if stage == 1 else (s2(S) if stage == 2 else s3(S))
    walk = hf(L, np.zeros(L.npts))
    orc.assign(walk, expr)
    orc.fill_periodic(walk)
    closed = hf(L, np.zeros(L.npts))
    orc.transport_stage(closed, PN, PN if stage == 1 else S, K, us, c, dt, stage)
    assert np.array_equal(closed.host, walk.host)


@pytest.mark.parametrize("threads", [2, 3, 8])
def test_threaded_stage_equals_serial(problem, threads):
    """The z-slab split of the stage (used at 512^3) gives the serial stage's bits,
    ghosts included, also with more threads than planes."""
    n, L, Ls, ops, k, phi, u, us, c, dt = problem
    s = periodic_random(L, np.random.default_rng(threads), lambda r, sh: r.standard_normal(sh))
    for stage in (1, 2, 3):
        a, b = hf(L, np.zeros(L.npts)), hf(L, np.zeros(L.npts))
        orc.transport_stage(a, hf(L, phi), hf(L, s), hf(L, k), us, c, dt, stage)
        orc.transport_stage(b, hf(L, phi), hf(L, s), hf(L, k), us, c, dt, stage, threads=threads)
        assert np.array_equal(a.host, b.host), stage


def _inner(L, a, b):
    return float(np.dot(interior(L, a).ravel(), interior(L, b).ravel()))


@pytest.mark.parametrize("stage", [1, 2, 3])
def test_adjoint_stage_dot_product_identity(problem, stage):
    n, L, Ls, ops, k, phi, u, us, c, dt = problem
    rng = np.random.default_rng(10 + stage)
    dt = 0.3 / n      # a larger step makes R's share of the identity visible
    gauss = lambda r, sh: r.standard_normal(sh)
    pn = periodic_random(L, rng, gauss)
    s = pn if stage == 1 else periodic_random(L, rng, gauss)
    a = periodic_random(L, rng, gauss)
    fwd = hf(L, np.zeros(L.npts))
    orc.transport_stage(fwd, hf(L, pn), hf(L, s), hf(L, k), us, c, dt, stage)
    out = hf(L, np.zeros(L.npts))
    acc_out = np.zeros(L.npts)
    acc_in = np.zeros(L.npts)
    gk = np.zeros(L.npts)
    orc.transport_adjoint_stage(out, acc_out, gk, hf(L, a), acc_in, hf(L, s), hf(L, k), us, c, dt, stage)
    lhs = _inner(L, a, fwd.host)
    if stage == 1:
        rhs = _inner(L, out.host, pn)
    else:
        rhs = _inner(L, out.host, s) + _inner(L, acc_out, pn)
    scale = np.abs(interior(L, a)).sum() * np.abs(interior(L, fwd.host)).max()
    assert abs(lhs - rhs) <= 1e-13 * scale
    # the adjoint output carries periodic ghosts
    g = out.host.reshape(L.glob[::-1])
    assert np.array_equal(g[0, 1:-1, 1:-1], g[-2, 1:-1, 1:-1])
    assert np.array_equal(g[1:-1, 1:-1, -1], g[1:-1, 1:-1, 1])


@pytest.mark.parametrize("stage", [1, 2, 3])
def test_adjoint_stage_k_part_is_exact_derivative(problem, stage):
    """The stage is affine in k: <a, F(k + dk) - F(k)> == <gk, dk> up to rounding."""
    n, L, Ls, ops, k, phi, u, us, c, dt = problem
    rng = np.random.default_rng(20 + stage)
    dt = 0.3 / n
    gauss = lambda r, sh: r.standard_normal(sh)
    pn = periodic_random(L, rng, gauss)
    s = pn if stage == 1 else periodic_random(L, rng, gauss)
    a = periodic_random(L, rng, gauss)
    dk = periodic_random(L, rng, gauss)
    f0, f1 = hf(L, np.zeros(L.npts)), hf(L, np.zeros(L.npts))
    orc.transport_stage(f0, hf(L, pn), hf(L, s), hf(L, k), us, c, dt, stage)
    orc.transport_stage(f1, hf(L, pn), hf(L, s), hf(L, k + dk), us, c, dt, stage)
    out = hf(L, np.zeros(L.npts))
    gk = np.zeros(L.npts)
    orc.transport_adjoint_stage(out, np.zeros(L.npts), gk, hf(L, a), np.zeros(L.npts), hf(L, s), hf(L, k),
                                us, c, dt, stage)
    lhs = _inner(L, a, f1.host - f0.host)
    rhs = _inner(L, gk, dk)
    assert abs(lhs) > 1e-6
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


def test_reverse_sweep_matches_central_differences():
    n, steps = 6, 3
    L, Ls, ops, k, phi, u, dt = W.transport_problem(n)
    dt = 0.2 / n
    phi = inputs.periodic_fill(phi.reshape(L.glob[::-1]).copy(), 1).reshape(-1)
    us = [nh.Field(l, host=x) for l, x in zip(Ls, u)]
    c = nh.axis_coefs(nh.build_stencils(n, n, n, 1.0, 1.0, 1.0))
    rng = np.random.default_rng(3)
    k = k * 200.0   # diffusion comparable to convection at this size
    # target: the trajectory of a perturbed k, so the loss is small
    kt = periodic_random(L, rng, lambda r, sh: interior(L, k) * (1.0 + 0.1 * r.standard_normal(sh)))
    _, target = orc.transport_forward_cpu(L, phi, hf(L, kt), us, c, dt, steps)

    def loss(kk, p0):
        _, pT = orc.transport_forward_cpu(L, p0, hf(L, kk), us, c, dt, steps)
        d = interior(L, pT - target)
        return 0.5 * float((d * d).sum())

    states, pT = orc.transport_forward_cpu(L, phi, hf(L, k), us, c, dt, steps)
    lamT = inputs.periodic_fill((pT - target).reshape(L.glob[::-1]).copy(), 1).reshape(-1)
    g0, gk = orc.transport_backward_cpu(L, lamT, hf(L, k), us, c, dt, states)
    eps = 1e-6
    for j in range(4):
        v = periodic_random(L, rng, lambda r, sh: r.standard_normal(sh))
        fd_k = (loss(k + eps * v, phi) - loss(k - eps * v, phi)) / (2 * eps)
        an_k = float((interior(L, gk) * interior(L, v)).sum())
        assert abs(fd_k - an_k) <= 1e-6 * max(abs(an_k), 1e-3), (j, fd_k, an_k)
        fd_p = (loss(k, phi + eps * v) - loss(k, phi - eps * v)) / (2 * eps)
        an_p = float((interior(L, g0) * interior(L, v)).sum())
        assert abs(fd_p - an_p) <= 1e-6 * max(abs(an_p), 1e-3), (j, fd_p, an_p)
