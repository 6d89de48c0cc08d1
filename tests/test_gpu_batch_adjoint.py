"""GPU tests of the batched persistent kernel (many uneven patches per launch,
hipGraph replay, autotuned tiles) and of the adjoint diffusion kernels with
checkpoint / recompute, against the oracle and finite differences."""
import math

import numpy as np
import pytest
import torch

import nebo_hip as nh
from nebo_hip import inputs
from nebo_hip import adjoint as adj
import oracle_lib as orc

pytestmark = pytest.mark.gpu


def _patch_dims(count, seed):
    sizes = np.array([16, 24, 32, 48, 64])
    idx = (inputs.splitmix64(seed, 3 * count) % np.uint64(5)).astype(np.int64).reshape(count, 3)
    return [tuple(int(v) for v in sizes[i]) for i in idx]


def _oracle_patch_step(d, a_host, coefs, s):
    L = nh.Layout.build(d, 1)
    src = nh.Field.from_numpy(L, a_host, where="cpu")
    dst = nh.Field.from_numpy(L, a_host.copy(), where="cpu")   # ghosts carried over
    orc.lapl_euler(dst, src, coefs, s)
    return dst.host


@pytest.mark.parametrize("count", [1, 7, 40])
def test_batch_matches_oracle_and_per_patch(count):
    dims = _patch_dims(count, 4 + count)
    db = nh.build_stencils(32, 32, 32, 1.0, 1.0, 1.0)
    coefs = nh.axis_coefs(db)
    s = 0.1 / (32.0 ** 2)
    rng = np.random.default_rng(count)
    init = [rng.standard_normal((d[0] + 2) * (d[1] + 2) * (d[2] + 2)) for d in dims]
    # the default tile, a 2-D tile run one launch per patch, flattened-plane
    # tiles: chunk lengths without (5) and with a compiled full-chunk path (4, 8,
    # 16; patch sides 16..64 also leave cut chunks), and column tiles (x = 0: 1, 2
    # or 4 rows per lane; items of up to 4 and 8 planes load all their planes as
    # they start)
    tiles = [None, (16, 8, 4), (128, 0, 5), (64, 0, 4), (256, 0, 8), (1024, 0, 16), (0, 1, 16), (0, 2, 5),
             (0, 4, 64), (0, 2, 4), (0, 4, 8), (0, 1, 3)]
    blocks = []
    for _ in tiles:
        a, b = nh.alloc_patches(dims)
        for x, h in zip(a, init):
            x.copy_(torch.from_numpy(h))
        for x, h in zip(b, init):  # ghost values of the b blocks = those of a
            x.copy_(torch.from_numpy(h))
        blocks.append((a, b))
    batches = [nh.Batch(a, b, dims, coefs, s, tile=t) for (a, b), t in zip(blocks, tiles)]
    for step in range(4):
        for j, B in enumerate(batches):
            (B.step_per_patch if j == 1 else B.step)(swap=step & 1)
    torch.cuda.synchronize()
    host = [h.copy() for h in init]
    for step in range(4):
        host = [_oracle_patch_step(d, h, coefs, s) for d, h in zip(dims, host)]
    for t, (a, _) in zip(tiles, blocks):
        for i in range(count):
            # 4 steps end in block a
            assert np.array_equal(a[i].cpu().numpy(), host[i]), f"tile {t} patch {i} dims {dims[i]}"
    for B in batches:
        B.close()


@pytest.mark.parametrize("tile", [(0, 1, 7), (0, 2, 16), (0, 4, 3), (0, 2, 4), (0, 4, 8), (0, 2, 6)])
def test_batch_column_tiles_odd_patches(tile):
    """Column tiles on patches narrower than a wave with widths that leave lanes
    idle (7, 24, 63), two columns wide (32 row groups per wave), exactly one wave
    wide, and wider than a wave (65 and 130 columns: strips of 64 and a last strip
    of 1 or 2), in one launch and one launch per patch: every patch equals the
    oracle's step."""
    dims = [(7, 9, 11), (2, 2, 2), (24, 5, 2), (63, 2, 7), (64, 3, 4), (65, 3, 5), (130, 5, 3), (2, 17, 3)]
    db = nh.build_stencils(32, 32, 32, 1.0, 1.0, 1.0)
    coefs = nh.axis_coefs(db)
    s = 0.1 / (32.0 ** 2)
    rng = np.random.default_rng(sum(tile))
    init = [rng.standard_normal((d[0] + 2) * (d[1] + 2) * (d[2] + 2)) for d in dims]
    want = [_oracle_patch_step(d, h, coefs, s) for d, h in zip(dims, init)]
    for per_patch in (False, True):
        a, b = nh.alloc_patches(dims)
        for blocks in (a, b):
            for x, h in zip(blocks, init):
                x.copy_(torch.from_numpy(h))
        B = nh.Batch(a, b, dims, coefs, s, tile=tile)
        (B.step_per_patch if per_patch else B.step)(swap=0)
        torch.cuda.synchronize()
        for i, d in enumerate(dims):
            assert np.array_equal(b[i].cpu().numpy(), want[i]), (tile, per_patch, d)
        B.close()


def test_batch_graph_replays_and_autotune():
    dims = _patch_dims(64, 4)
    db = nh.build_stencils(32, 32, 32, 1.0, 1.0, 1.0)
    coefs = nh.axis_coefs(db)
    s = 0.05 / (32.0 ** 2)
    rng = np.random.default_rng(1)
    init = [rng.standard_normal((d[0] + 2) * (d[1] + 2) * (d[2] + 2)) for d in dims]
    a, b = nh.alloc_patches(dims)
    ea, eb = nh.alloc_patches(dims)
    for blocks in (a, b, ea, eb):
        for x, h in zip(blocks, init):
            x.copy_(torch.from_numpy(h))
    B = nh.Batch(a, b, dims, coefs, s)
    tile = B.autotune()
    assert len(tile) == 3 and (tile[0] * tile[1] <= 1024 if tile[1] else tile[0] % 64 == 0)
    for x, h in zip(a, init):       # autotune stepped the blocks; restore
        x.copy_(torch.from_numpy(h))
    for x, h in zip(b, init):
        x.copy_(torch.from_numpy(h))
    E = nh.Batch(ea, eb, dims, coefs, s, tile=(64, 4, 16))
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        B.graph_build(10, stream)
        for rep in range(3):
            B.graph_launch(stream)
            stream.synchronize()
            for st in range(10):
                E.step(swap=st & 1, stream=stream)
            stream.synchronize()
            # eager work between replays, then compare every patch
            for i in range(len(dims)):
                assert torch.equal(a[i], ea[i]), f"replay {rep} patch {i}"
    B.close()
    E.close()


# --------------------------------------------------------------------------
# adjoint
# --------------------------------------------------------------------------

def _periodic(L, arr):
    a = np.zeros(L.glob[::-1])
    n = L.interior
    a[1:-1, 1:-1, 1:-1] = arr.reshape(n[2], n[1], n[0])
    return inputs.periodic_fill(a, 1).reshape(-1)


@pytest.mark.parametrize("n", [(16, 16, 16), (13, 10, 9)])
def test_diffusion_step_and_adjoint_bit_exact(n):
    L = nh.Layout.build(n, 1)
    db = nh.build_stencils(*n, 1.0, 1.0, 1.0)
    c = nh.axis_coefs(db)
    rng = np.random.default_rng(sum(n))
    N = n[0] * n[1] * n[2]
    k = _periodic(L, rng.uniform(0.5, 2.0, N))
    phi = _periodic(L, rng.standard_normal(N))
    lam = _periodic(L, rng.standard_normal(N))
    dt = 0.1 / (max(n) ** 2 * 2.0)
    g = {name: torch.from_numpy(v).cuda() for name, v in (("k", k), ("phi", phi), ("lam", lam))}
    out = torch.zeros_like(g["phi"])
    adj.diffusion_step(L, out, g["phi"], g["k"], c, dt)
    lam_out = torch.zeros_like(g["phi"])
    gk = torch.zeros_like(g["phi"])
    adj.diffusion_adjoint_step(L, lam_out, gk, g["lam"], g["phi"], g["k"], c, dt)
    torch.cuda.synchronize()
    ro = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.diffusion_step(ro, nh.Field(L, host=phi), nh.Field(L, host=k), c, dt)
    rl = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    rg = np.zeros(L.npts)
    orc.diffusion_adjoint_step(rl, rg, nh.Field(L, host=lam), nh.Field(L, host=phi), nh.Field(L, host=k), c, dt)
    assert np.array_equal(out.cpu().numpy(), ro.host)
    assert np.array_equal(lam_out.cpu().numpy(), rl.host)
    assert np.array_equal(gk.cpu().numpy(), rg)


def test_diffusion_step_every_tile_bit_exact():
    """The Euler form of the diffusion kernel under every stencil tile."""
    n = (70, 21, 29)
    L = nh.Layout.build(n, 1)
    c = nh.axis_coefs(nh.build_stencils(*n, 1.0, 1.0, 1.0))
    rng = np.random.default_rng(77)
    N = n[0] * n[1] * n[2]
    k = _periodic(L, rng.uniform(0.5, 2.0, N))
    phi = _periodic(L, rng.standard_normal(N))
    dt = 1e-3
    ro = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    orc.diffusion_step(ro, nh.Field(L, host=phi), nh.Field(L, host=k), c, dt)
    gp, gkk = torch.from_numpy(phi).cuda(), torch.from_numpy(k).cuda()
    try:
        for tile in nh.stencil_tiles():
            nh.stencil_set_tile(0, *tile)
            out = torch.zeros_like(gp)
            adj.diffusion_step(L, out, gp, gkk, c, dt)
            torch.cuda.synchronize()
            assert np.array_equal(out.cpu().numpy(), ro.host), tile
    finally:
        nh.stencil_reset_tile(0)


def _trajectory_case(n, steps, seed=5):
    L = nh.Layout.build((n, n, n), 1)
    db = nh.build_stencils(n, n, n, 1.0, 1.0, 1.0)
    c = nh.axis_coefs(db)
    N = n ** 3
    rng = np.random.default_rng(seed)
    k = _periodic(L, inputs.uniform(2, N, 0.5, 2.0))
    # the periodic step works on the wrapped interior: give phi0 exact periodic ghosts
    phi0 = inputs.periodic_fill(inputs.fourier_field(n, 3).copy(), 1).reshape(-1)
    dt = 0.1 * (1.0 / n) ** 2 / 2.0
    return L, c, k, phi0, dt, rng


def _smoothed_perturbation(L, k, rng):
    """k_true = k*(1 + 0.1*N(0,1)) smoothed by three interpolant passes (cell -> face ->
    cell, one per axis, periodic)."""
    n = L.interior
    a = k.reshape(L.glob[::-1])[1:-1, 1:-1, 1:-1] * (1.0 + 0.1 * rng.standard_normal((n[2], n[1], n[0])))
    for ax in (2, 1, 0):       # numpy axes z, y, x -> x, y, z passes
        face = 0.5 * np.roll(a, 1, axis=ax) + 0.5 * a            # low face of each cell
        a = 0.5 * face + 0.5 * np.roll(face, -1, axis=ax)        # back to cells
    return _periodic(L, a.reshape(-1))


def test_checkpointed_trajectory_bit_exact_vs_oracle():
    n, steps, every = 12, 8, 4
    L, c, k, phi0, dt, rng = _trajectory_case(n, steps)
    target = _periodic(L, rng.standard_normal(n ** 3))
    gk_, gp_ = torch.from_numpy(k).cuda(), torch.from_numpy(phi0).cuda()
    phiT, ckpt = adj.forward(L, gp_, gk_, c, dt, steps, every)
    lamT = phiT - torch.from_numpy(target).cuda()
    lam0, gk = adj.backward(L, lamT, gk_, c, dt, steps, every, ckpt)
    torch.cuda.synchronize()
    # oracle: plain loop keeping every state
    states = [phi0.copy()]
    for _ in range(steps):
        o = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
        orc.diffusion_step(o, nh.Field(L, host=states[-1]), nh.Field(L, host=k), c, dt)
        states.append(o.host)
    assert np.array_equal(phiT.cpu().numpy(), states[-1])
    lam = states[-1] - target
    lam = inputs.periodic_fill(lam.reshape(L.glob[::-1]).copy(), 1).reshape(-1)
    g = np.zeros(L.npts)
    for s in range(steps - 1, -1, -1):
        o = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
        orc.diffusion_adjoint_step(o, g, nh.Field(L, host=lam), nh.Field(L, host=states[s]),
                                   nh.Field(L, host=k), c, dt)
        lam = o.host
    assert np.array_equal(lam0.cpu().numpy(), lam)
    assert np.array_equal(gk.cpu().numpy(), g)


def test_autograd_gradient_matches_finite_differences():
    """Adjoint: dL/dk and dL/dphi0 within 1e-6 (relative) of central differences
    (eps = 1e-6) on random directions; 32^3, 20 steps, checkpoint every 5.  The
    target is the trajectory of a perturbed, smoothed k, so the loss is small and
    the differences are not swamped by rounding of the loss itself."""
    n, steps, every = 32, 20, 5
    L, c, k, phi0, dt, rng = _trajectory_case(n, steps)
    k_true = torch.from_numpy(_smoothed_perturbation(L, k, rng)).cuda()
    target, _ = adj.forward(L, torch.from_numpy(phi0).cuda(), k_true, c, dt, steps, every)
    inner = torch.zeros(L.glob[::-1], dtype=torch.bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.reshape(-1).cuda()

    def loss_of(kk, pp):
        out = adj.DiffusionTrajectory.apply(pp, kk, L, c, dt, steps, every)
        d = (out - target)[inner]
        return 0.5 * (d * d).sum()

    kt = torch.from_numpy(k).cuda().requires_grad_(True)
    pt = torch.from_numpy(phi0).cuda().requires_grad_(True)
    loss = loss_of(kt, pt)
    loss.backward()
    gk, gp = kt.grad.detach(), pt.grad.detach()
    eps = 1e-6
    with torch.no_grad():
        for j in range(6):
            v = torch.from_numpy(_periodic(L, rng.standard_normal(n ** 3))).cuda()
            for which, grad in (("k", gk), ("phi", gp)):
                if which == "k":
                    lp, lm = loss_of(kt + eps * v, pt), loss_of(kt - eps * v, pt)
                else:
                    lp, lm = loss_of(kt, pt + eps * v), loss_of(kt, pt - eps * v)
                fd = (lp - lm).item() / (2 * eps)
                an = (grad[inner] * v[inner]).sum().item()
                assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (which, j, fd, an)
