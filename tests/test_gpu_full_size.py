"""The benchmark workloads at their stated sizes on the GPU, checked against the
oracle (each test prints its case before the GPU work starts).

* 512 uneven patches (interior sides drawn from {16, 24, 32, 48, 64} per axis,
  seeded), Laplacian + Euler: 10 steps replayed from one hipGraph of persistent
  launches, and 10 steps of one launch per patch (the per-patch launch semantics
  of spatialops/NeboLhs.h:400-478), each patch equal to the oracle's walk.
* Diffusion adjoint, 128^3, 100 explicit Euler steps, checkpoint every 10: the
  forward trajectory, dL/dphi0 and dL/dk bit-exact with the oracle's CPU adjoint,
  and central differences (eps = 1e-6) on 32 seeded directions for k and for phi0.
* Scalar transport, 512^3: one SSP-RK3 step (three stages) of the hand-written
  stage kernel and of the fused generic assignments, bit-exact with the oracle's
  stages (z-slab threaded); 20 steps conserving the interior sum; the 20-step
  adjoint with checkpoints every 5 steps bit-exact with the oracle at 64^3 and
  within 1e-6 of central differences at 512^3.
"""
import os

import numpy as np
import pytest
import torch

import nebo_hip as nh
from nebo_hip import inputs
from nebo_hip import adjoint as adj
from nebo_hip import transport as tr
from nebo_hip import workloads as W
import oracle_lib as orc

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _say(msg):
    print(msg, flush=True)


# --------------------------------------------------------------------------
# 512 uneven patches
# --------------------------------------------------------------------------

def _patch_dims(count, seed):
    sizes = np.array([16, 24, 32, 48, 64])
    idx = (inputs.splitmix64(seed, 3 * count) % np.uint64(5)).astype(np.int64).reshape(count, 3)
    return [tuple(int(v) for v in sizes[i]) for i in idx]


def test_512_patches_graph_and_per_patch_match_oracle():
    count, steps = 512, 10
    _say(f"{count} patches, {steps}-step hipGraph and per-patch launches")
    dims = _patch_dims(count, 4)
    coefs = nh.axis_coefs(nh.build_stencils(64, 64, 64, 1.0, 1.0, 1.0))
    s = 0.1 / 64.0 ** 2
    rng = np.random.default_rng(512)
    init = [rng.standard_normal((d[0] + 2) * (d[1] + 2) * (d[2] + 2)) for d in dims]
    a, b = nh.alloc_patches(dims)
    pa, pb = nh.alloc_patches(dims)
    for blocks in (a, b, pa, pb):
        for x, h in zip(blocks, init):
            x.copy_(torch.from_numpy(h))
    B = nh.Batch(a, b, dims, coefs, s)
    tile = B.autotune()            # the tile the benchmark runs; it stepped a -> b
    for blocks in (a, b):
        for x, h in zip(blocks, init):
            x.copy_(torch.from_numpy(h))
    P = nh.Batch(pa, pb, dims, coefs, s, tile=tile)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        B.graph_build(steps, stream)
        B.graph_launch(stream)
        for st in range(steps):
            P.step_per_patch(swap=st & 1, stream=stream)
        stream.synchronize()
    # oracle: every patch walked on its own, ghost values carried over
    for i, (d, h) in enumerate(zip(dims, init)):
        L = nh.Layout.build(d, 1)
        cur = h
        for _ in range(steps):
            dst = nh.Field.from_numpy(L, cur.copy(), where="cpu")
            orc.lapl_euler(dst, nh.Field(L, host=cur), coefs, s)
            cur = dst.host
        # an even number of steps ends in block a
        assert np.array_equal(a[i].cpu().numpy(), cur), f"graph, tile {tile}, patch {i} dims {d}"
        assert np.array_equal(pa[i].cpu().numpy(), cur), f"per patch, patch {i} dims {d}"
    B.close()
    P.close()


# --------------------------------------------------------------------------
# diffusion adjoint, 128^3 x 100 steps
# --------------------------------------------------------------------------

def _periodic(L, arr):
    a = np.zeros(L.glob[::-1])
    n = L.interior
    a[1:-1, 1:-1, 1:-1] = arr.reshape(n[2], n[1], n[0])
    return inputs.periodic_fill(a, 1).reshape(-1)


def _smoothed_perturbation(L, k, rng):
    """k_true = k*(1 + 0.1*N(0,1)) smoothed by three interpolant passes (cell ->
    face -> cell, one per axis, periodic)."""
    n = L.interior
    a = k.reshape(L.glob[::-1])[1:-1, 1:-1, 1:-1] * (1.0 + 0.1 * rng.standard_normal((n[2], n[1], n[0])))
    for ax in (2, 1, 0):
        face = 0.5 * np.roll(a, 1, axis=ax) + 0.5 * a
        a = 0.5 * face + 0.5 * np.roll(face, -1, axis=ax)
    return _periodic(L, a.reshape(-1))


@pytest.fixture(scope="module")
def diffusion_128():
    n, steps, every = 128, 100, 10
    L = nh.Layout.build((n, n, n), 1)
    c = nh.axis_coefs(nh.build_stencils(n, n, n, 1.0, 1.0, 1.0))
    k = inputs.periodic_fill(inputs.coefficient_field(n, 5).copy(), 1).reshape(-1)
    phi0 = inputs.periodic_fill(inputs.fourier_field(n, 3).copy(), 1).reshape(-1)
    dt = 0.1 * (1.0 / n) ** 2 / float(k.max())
    k_true = _smoothed_perturbation(L, k, np.random.default_rng(5))
    return dict(n=n, steps=steps, every=every, L=L, c=c, k=k, phi0=phi0, dt=dt, k_true=k_true)


def _oracle_trajectory(P, k):
    L = P["L"]
    states = [P["phi0"]]
    for _ in range(P["steps"]):
        o = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
        orc.diffusion_step(o, nh.Field(L, host=states[-1]), nh.Field(L, host=k), P["c"], P["dt"])
        states.append(o.host)
    return states


def test_adjoint_128_100_steps_bit_exact_vs_oracle(diffusion_128):
    P = diffusion_128
    L, c, dt, steps, every = P["L"], P["c"], P["dt"], P["steps"], P["every"]
    _say(f"diffusion adjoint {P['n']}^3, {steps} steps, checkpoint every {every}")
    gk_ = torch.from_numpy(P["k"]).cuda()
    gp_ = torch.from_numpy(P["phi0"]).cuda()
    target, _ = adj.forward(L, gp_, torch.from_numpy(P["k_true"]).cuda(), c, dt, steps, every)
    phiT, ckpt = adj.forward(L, gp_, gk_, c, dt, steps, every)
    lamT = phiT - target
    lam0, gk = adj.backward(L, lamT, gk_, c, dt, steps, every, ckpt)
    torch.cuda.synchronize()
    tgt = target.cpu().numpy()
    assert np.array_equal(tgt, _oracle_trajectory(P, P["k_true"])[-1]), "target trajectory"
    states = _oracle_trajectory(P, P["k"])
    assert np.array_equal(phiT.cpu().numpy(), states[-1]), "forward trajectory"
    lam = inputs.periodic_fill((states[-1] - tgt).reshape(L.glob[::-1]).copy(), 1).reshape(-1)
    g = np.zeros(L.npts)
    for s in range(steps - 1, -1, -1):
        o = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
        orc.diffusion_adjoint_step(o, g, nh.Field(L, host=lam), nh.Field(L, host=states[s]),
                                   nh.Field(L, host=P["k"]), c, dt)
        lam = o.host
    assert np.array_equal(lam0.cpu().numpy(), lam), "dL/dphi0"
    assert np.array_equal(gk.cpu().numpy(), g), "dL/dk"


def test_adjoint_128_32_directions_vs_finite_differences(diffusion_128):
    """L = 1/2 |phi_T - phi_target|^2 over the interior; central differences with
    eps = 1e-6 on 32 seeded directions for k and 32 for phi0, each within 1e-6
    (relative to max(1, |directional derivative|)) of the adjoint's."""
    P = diffusion_128
    L, c, dt, steps, every = P["L"], P["c"], P["dt"], P["steps"], P["every"]
    _say(f"diffusion adjoint {P['n']}^3 x {steps}: 32 FD directions")
    target, _ = adj.forward(L, torch.from_numpy(P["phi0"]).cuda(), torch.from_numpy(P["k_true"]).cuda(), c, dt,
                            steps, every)
    inner = torch.zeros(L.glob[::-1], dtype=torch.bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.reshape(-1).cuda()

    def loss_of(kk, pp):
        out = adj.DiffusionTrajectory.apply(pp, kk, L, c, dt, steps, every)
        d = (out - target)[inner]
        return 0.5 * (d * d).sum()

    kt = torch.from_numpy(P["k"]).cuda().requires_grad_(True)
    pt = torch.from_numpy(P["phi0"]).cuda().requires_grad_(True)
    loss_of(kt, pt).backward()
    gk, gp = kt.grad.detach(), pt.grad.detach()
    eps = 1e-6
    n3 = P["n"] ** 3
    worst = 0.0
    with torch.no_grad():
        for j in range(32):
            for which, grad in (("k", gk), ("phi0", gp)):
                v = torch.from_numpy(_periodic(L, inputs.normal(6000 + 2 * j + (which == "k"), n3))).cuda()
                if which == "k":
                    lp, lm = loss_of(kt + eps * v, pt), loss_of(kt - eps * v, pt)
                else:
                    lp, lm = loss_of(kt, pt + eps * v), loss_of(kt, pt - eps * v)
                fd = (lp - lm).item() / (2 * eps)
                an = (grad[inner] * v[inner]).sum().item()
                err = abs(fd - an) / max(1.0, abs(an))
                worst = max(worst, err)
                assert err <= 1e-6, (which, j, fd, an)
    _say(f"worst relative FD difference {worst:.3e}")


# --------------------------------------------------------------------------
# scalar transport, 512^3
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def transport_512():
    n = 512
    L, Ls, ops, k, phi, u, dt = W.transport_problem_gpu(n)
    c = nh.axis_coefs(nh.build_stencils(n, n, n, 1.0, 1.0, 1.0))
    torch.cuda.synchronize()
    return dict(n=n, L=L, Ls=Ls, ops=ops, k=k, phi=phi, u=u, dt=dt, c=c)


def test_transport_512_one_step_bit_exact_vs_oracle(transport_512):
    """The inputs are generated on the GPU and copied to the host, so both sides
    read the same bits.  Hand-written stage kernel (the schedule the box-size rule
    picks at 512^3: column tiles, 512-plane chunks) and the three fused stage
    assignments of the generic evaluator, each stage against the oracle's."""
    T = transport_512
    L, Ls, k, phi, u, dt, c = T["L"], T["Ls"], T["k"], T["phi"], T["u"], T["dt"], T["c"]
    _say(f"transport {T['n']}^3, one SSP-RK3 step, schedule {tr.schedule(L)}")
    hk, hphi = k.cpu().numpy(), phi.cpu().numpy()
    Uh = [nh.Field(l, host=x.cpu().numpy()) for l, x in zip(Ls, u)]
    hand = [torch.empty_like(phi) for _ in range(3)]
    tr.stage(L, hand[0], phi, None, k, u, c, dt, 1, face_layouts=Ls)
    tr.stage(L, hand[1], phi, hand[0], k, u, c, dt, 2, face_layouts=Ls)
    tr.stage(L, hand[2], phi, hand[1], k, u, c, dt, 3, face_layouts=Ls)
    K, PN = nh.Field(L, gpu=k), nh.Field(L, gpu=phi)
    U = [nh.Field(l, gpu=x) for l, x in zip(Ls, u)]
    G = [nh.Field(L, gpu=torch.empty_like(phi)) for _ in range(3)]
    lo, hi = W.upwind_ops()
    s1, s2, s3 = W.rk3_stage_exprs(T["ops"], PN, K, U, dt, lo, hi)
    for g, e in zip(G, (s1(), s2(G[0]), s3(G[1]))):
        nh.assign(g, e)
        nh.fill_ghosts_periodic(g)
    torch.cuda.synchronize()
    prev = hphi
    for stage in (1, 2, 3):
        ref = nh.Field(L, host=np.zeros(L.npts))
        orc.transport_stage(ref, nh.Field(L, host=hphi), nh.Field(L, host=prev), nh.Field(L, host=hk), Uh, c,
                            dt, stage, threads=THREADS)
        got = hand[stage - 1].cpu().numpy()
        assert np.array_equal(got, ref.host), f"stage {stage}, hand-written kernel"
        assert np.array_equal(G[stage - 1].gpu.cpu().numpy(), ref.host), f"stage {stage}, generic evaluator"
        prev = ref.host
        del got


def test_transport_512_twenty_steps_conserve(transport_512):
    """20 SSP-RK3 steps: the divergence-free upwind flux and the diffusive flux are
    conservative on the periodic box, so the interior sum changes only by
    rounding (relative to the sum of |phi|)."""
    T = transport_512
    L, Ls, k, phi, u, dt, c = T["L"], T["Ls"], T["k"], T["phi"], T["u"], T["dt"], T["c"]
    _say(f"transport {T['n']}^3, 20 steps, conservation")
    phiT, _ = tr.forward(L, phi, k, u, c, dt, 20, 5, face_layouts=Ls)
    torch.cuda.synchronize()
    n = T["n"]
    inner0 = phi.view(n + 2, n + 2, n + 2)[1:-1, 1:-1, 1:-1]
    inner1 = phiT.view(n + 2, n + 2, n + 2)[1:-1, 1:-1, 1:-1]
    s0, s1 = inner0.sum().item(), inner1.sum().item()
    a0 = inner0.abs().sum().item()
    assert torch.isfinite(inner1).all()
    assert abs(s1 - s0) <= 1e-12 * a0, (s0, s1, a0)
    assert not torch.equal(inner0, inner1)


def test_transport_adjoint_20_steps_bit_exact_vs_oracle_64():
    """The 20-step adjoint with checkpoints every 5 steps, at a size the oracle's
    reverse sweep covers in seconds: phi_T, dL/dphi0 and dL/dk bit-exact."""
    n, steps, every = 64, 20, 5
    _say(f"transport adjoint {n}^3, {steps} steps, checkpoint every {every}")
    L, Ls, ops, k, phi, u, dt = W.transport_problem(n)
    phi = inputs.periodic_fill(phi.reshape(L.glob[::-1]).copy(), 1).reshape(-1)
    c = nh.axis_coefs(nh.build_stencils(n, n, n, 1.0, 1.0, 1.0))
    rng = np.random.default_rng(13)
    Uh = [nh.Field(l, host=x) for l, x in zip(Ls, u)]
    Ug = [torch.from_numpy(x).cuda() for x in u]
    gk_, gp_ = torch.from_numpy(k).cuda(), torch.from_numpy(phi).cuda()
    phiT, ckpt = tr.forward(L, gp_, gk_, Ug, c, dt, steps, every, face_layouts=Ls)
    target = _periodic(L, rng.standard_normal(n ** 3))
    lamT = phiT - torch.from_numpy(target).cuda()
    lam0, gk = tr.backward(L, lamT, gk_, Ug, c, dt, steps, every, ckpt, face_layouts=Ls)
    torch.cuda.synchronize()
    states, pT = orc.transport_forward_cpu(L, phi, nh.Field(L, host=k), Uh, c, dt, steps)
    assert np.array_equal(phiT.cpu().numpy(), pT)
    lam = inputs.periodic_fill((pT - target).reshape(L.glob[::-1]).copy(), 1).reshape(-1)
    g0, gkh = orc.transport_backward_cpu(L, lam, nh.Field(L, host=k), Uh, c, dt, states)
    assert np.array_equal(lam0.cpu().numpy(), g0)
    assert np.array_equal(gk.cpu().numpy(), gkh)


def test_transport_512_adjoint_vs_finite_differences(transport_512):
    """dL/dk and dL/dphi0 of L = 1/2 |phi_T - t|^2 after 20 steps at 512^3 (checkpoint
    every 5), t the trajectory of a perturbed k: two seeded directions each, within
    1e-6 of central differences (eps = 1e-6).  k is about 1e-3 here, so its
    directions are relative (k times N(0,1) per cell): eps*v then changes k by
    1e-6 of itself, and the O(eps^2) error of the central difference stays far
    below the tolerance (a plain N(0,1) direction would change k by 1e-3 of
    itself)."""
    T = transport_512
    L, Ls, k, phi, u, dt, c = T["L"], T["Ls"], T["k"], T["phi"], T["u"], T["dt"], T["c"]
    n, steps, every = T["n"], 20, 5
    _say(f"transport adjoint {n}^3, {steps} steps, checkpoint every {every}: FD directions")
    gen = torch.Generator(device="cuda").manual_seed(17)

    def periodic_dir():
        a = torch.zeros((n + 2,) * 3, dtype=torch.float64, device="cuda")
        a[1:-1, 1:-1, 1:-1] = torch.randn((n, n, n), generator=gen, device="cuda", dtype=torch.float64)
        a[0], a[-1] = a[-2].clone(), a[1].clone()
        a[:, 0], a[:, -1] = a[:, -2].clone(), a[:, 1].clone()
        a[:, :, 0], a[:, :, -1] = a[:, :, -2].clone(), a[:, :, 1].clone()
        return a.reshape(-1)

    k_true = k * (1.0 + 0.1 * periodic_dir())
    target, _ = tr.forward(L, phi, k_true, u, c, dt, steps, every, face_layouts=Ls)
    del k_true
    inner = torch.zeros((n + 2,) * 3, dtype=torch.bool, device="cuda")
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.reshape(-1)

    def loss_of(kk, pp):
        out = tr.TransportTrajectory.apply(pp, kk, L, Ls, u, c, dt, steps, every)
        d = (out - target)[inner]
        return 0.5 * (d * d).sum()

    kt = k.clone().requires_grad_(True)
    pt = phi.clone().requires_grad_(True)
    loss_of(kt, pt).backward()
    gk, gp = kt.grad.detach(), pt.grad.detach()
    eps = 1e-6
    with torch.no_grad():
        for j in range(2):
            for which, grad in (("k", gk), ("phi0", gp)):
                v = periodic_dir() * (k if which == "k" else 1.0)
                if which == "k":
                    lp, lm = loss_of(kt + eps * v, pt), loss_of(kt - eps * v, pt)
                else:
                    lp, lm = loss_of(kt, pt + eps * v), loss_of(kt, pt - eps * v)
                fd = (lp - lm).item() / (2 * eps)
                an = (grad[inner] * v[inner]).sum().item()
                _say(f"  {which} direction {j}: fd {fd:.12e} adjoint {an:.12e}")
                assert abs(fd - an) <= 1e-6 * max(1e-3, abs(an)), (which, j, fd, an)
