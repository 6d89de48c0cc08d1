"""The scalar-transport chain on the GPU (hand-written SSP-RK3 stage kernel, its
transpose, checkpointed forward / reverse sweeps) against the generic evaluator,
the oracle and central finite differences."""
import numpy as np
import pytest
import torch

import nebo_hip as nh
from nebo_hip import inputs
from nebo_hip import transport as tr
from nebo_hip import workloads as W
import oracle_lib as orc

pytestmark = pytest.mark.gpu


def _problem(n, k_scale=1.0):
    L, Ls, ops, k, phi, u, dt = W.transport_problem(n)
    phi = inputs.periodic_fill(phi.reshape(L.glob[::-1]).copy(), 1).reshape(-1)
    c = nh.axis_coefs(nh.build_stencils(n, n, n, 1.0, 1.0, 1.0))
    return L, Ls, ops, k * k_scale, phi, u, dt, c


def _periodic_random(L, rng):
    n = L.interior
    a = np.zeros(L.glob[::-1])
    a[1:-1, 1:-1, 1:-1] = rng.standard_normal((n[2], n[1], n[0]))
    return inputs.periodic_fill(a, 1).reshape(-1)


@pytest.fixture(params=[(0, 0), (1, 5), (1, 64), (2, 5), (2, 64), (3, 7), (3, 16)],
                ids=["auto", "direct-z5", "direct-z64", "column-z5", "column-z64", "column-pf1-z7",
                     "column-pf1-z16"])
def zchunk(request):
    """The stage kernels under each schedule and several z chunks (5 and 7 divide
    none of the boxes; 64 exceeds them)."""
    tr.tune(*request.param)
    yield request.param
    tr.tune()


def test_schedule_by_box_size():
    """Column tiles with the longest z chunk that fills every CU on large boxes,
    the direct schedule on small ones; a tuned schedule applies as set."""
    tr.tune()
    assert tr.schedule(nh.Layout.build((512, 512, 512), 1)) == (tr.SCHED_COLUMN_PF1, 512)
    assert tr.schedule(nh.Layout.build((256, 256, 256), 1)) == (tr.SCHED_COLUMN_PF1, 64)
    assert tr.schedule(nh.Layout.build((64, 64, 64), 1)) == (tr.SCHED_DIRECT, 16)
    try:
        tr.tune(tr.SCHED_COLUMN, 9)
        assert tr.schedule(nh.Layout.build((64, 64, 64), 1)) == (tr.SCHED_COLUMN, 9)
    finally:
        tr.tune()


@pytest.mark.parametrize("n", [20, 37, 72])
def test_stage_kernel_matches_generic_evaluator_and_oracle(n, zchunk):
    """Each stage on the hand-written kernel, the generic evaluator's compiled
    kernel and the oracle, bit for bit.  At 72 the compiled stages hold two
    columns per lane (boxes wider than 64), with the wave-edge values read by
    every lane outside the selects that use them."""
    L, Ls, ops, k, phi, u, dt, c = _problem(n)
    rng = np.random.default_rng(n)
    s = _periodic_random(L, rng)
    lo, hi = W.upwind_ops()
    Ug = [nh.Field.from_numpy(l, x, where="gpu") for l, x in zip(Ls, u)]
    Uh = [nh.Field(l, host=x) for l, x in zip(Ls, u)]
    K, PN, S = (nh.Field.from_numpy(L, a, where="both") for a in (k, phi, s))
    for stage in (1, 2, 3):
        s1, s2, s3 = W.rk3_stage_exprs(ops, PN, K, Ug, dt, lo, hi)
        expr = s1() if stage == 1 else (s2(S) if stage == 2 else s3(S))
        gen = nh.Field.empty(L, "gpu")
        gen.gpu.zero_()
        nh.assign(gen, expr)
        nh.fill_ghosts_periodic(gen)
        hand = nh.Field.empty(L, "gpu")
        hand.gpu.zero_()
        tr.stage(L, hand, PN, S if stage != 1 else None, K, Ug, c, dt, stage)
        ref = nh.Field(L, host=np.zeros(L.npts))
        orc.transport_stage(ref, nh.Field(L, host=phi), nh.Field(L, host=s if stage != 1 else phi),
                            nh.Field(L, host=k), Uh, c, dt, stage)
        torch.cuda.synchronize()
        got = hand.gpu.cpu().numpy()
        assert np.array_equal(got, gen.gpu.cpu().numpy()), f"stage {stage} vs generic evaluator"
        assert np.array_equal(got, ref.host), f"stage {stage} vs oracle"


@pytest.mark.parametrize("n", [16, 23])
def test_adjoint_stage_bit_exact(n, zchunk):
    L, Ls, ops, k, phi, u, dt, c = _problem(n)
    rng = np.random.default_rng(100 + n)
    a, s, acc_in = _periodic_random(L, rng), _periodic_random(L, rng), rng.standard_normal(L.npts)
    Uh = [nh.Field(l, host=x) for l, x in zip(Ls, u)]
    Ug = [torch.from_numpy(x).cuda() for x in u]
    g = {name: torch.from_numpy(v).cuda() for name, v in (("a", a), ("s", s), ("k", k), ("acc", acc_in))}
    for stage in (1, 2, 3):
        out = torch.zeros(L.npts, dtype=torch.float64, device="cuda")
        acc_out = torch.zeros_like(out)
        gk = torch.from_numpy(np.full(L.npts, 0.5)).cuda()
        tr.adjoint_stage(L, out, acc_out if stage != 1 else None, gk, g["a"], g["acc"] if stage != 3 else None,
                         g["s"], g["k"], Ug, c, dt, stage, face_layouts=Ls)
        ro = nh.Field(L, host=np.zeros(L.npts))
        racc = np.zeros(L.npts)
        rgk = np.full(L.npts, 0.5)
        orc.transport_adjoint_stage(ro, racc if stage != 1 else None, rgk, nh.Field(L, host=a),
                                    acc_in if stage != 3 else None, nh.Field(L, host=s), nh.Field(L, host=k), Uh,
                                    c, dt, stage)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ro.host), f"stage {stage} out"
        assert np.array_equal(gk.cpu().numpy(), rgk), f"stage {stage} gk"
        if stage != 1:
            got = acc_out.cpu().numpy().reshape(L.glob[::-1])[1:-1, 1:-1, 1:-1]
            assert np.array_equal(got, racc.reshape(L.glob[::-1])[1:-1, 1:-1, 1:-1]), f"stage {stage} acc"


def test_checkpointed_trajectory_bit_exact_vs_oracle(zchunk):
    n, steps, every = 12, 6, 3
    L, Ls, ops, k, phi, u, dt, c = _problem(n, k_scale=50.0)
    rng = np.random.default_rng(9)
    Uh = [nh.Field(l, host=x) for l, x in zip(Ls, u)]
    Ug = [torch.from_numpy(x).cuda() for x in u]
    gk_, gp_ = torch.from_numpy(k).cuda(), torch.from_numpy(phi).cuda()
    phiT, ckpt = tr.forward(L, gp_, gk_, Ug, c, dt, steps, every, face_layouts=Ls)
    target = _periodic_random(L, rng)
    lamT = phiT - torch.from_numpy(target).cuda()
    lam0, gk = tr.backward(L, lamT, gk_, Ug, c, dt, steps, every, ckpt, face_layouts=Ls)
    torch.cuda.synchronize()
    states, pT = orc.transport_forward_cpu(L, phi, nh.Field(L, host=k), Uh, c, dt, steps)
    assert np.array_equal(phiT.cpu().numpy(), pT)
    lam = inputs.periodic_fill((pT - target).reshape(L.glob[::-1]).copy(), 1).reshape(-1)
    g0, gkh = orc.transport_backward_cpu(L, lam, nh.Field(L, host=k), Uh, c, dt, states)
    assert np.array_equal(lam0.cpu().numpy(), g0)
    assert np.array_equal(gk.cpu().numpy(), gkh)


def test_autograd_gradient_matches_finite_differences():
    """dL/dk and dL/dphi0 of L = 1/2 |phi_T - t|^2 within 1e-6 of central differences
    (eps = 1e-6); 24^3, 10 RK3 steps, checkpoint every 5; t is the trajectory of a
    perturbed k."""
    n, steps, every = 24, 10, 5
    L, Ls, ops, k, phi, u, dt, c = _problem(n, k_scale=20.0)
    rng = np.random.default_rng(11)
    Ug = [torch.from_numpy(x).cuda() for x in u]
    inner = torch.zeros(L.glob[::-1], dtype=torch.bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.reshape(-1).cuda()
    kp = k.reshape(L.glob[::-1])[1:-1, 1:-1, 1:-1] * (1.0 + 0.1 * rng.standard_normal((n, n, n)))
    a = np.zeros(L.glob[::-1])
    a[1:-1, 1:-1, 1:-1] = kp
    k_true = torch.from_numpy(inputs.periodic_fill(a, 1).reshape(-1)).cuda()
    target, _ = tr.forward(L, torch.from_numpy(phi).cuda(), k_true, Ug, c, dt, steps, every, face_layouts=Ls)

    def loss_of(kk, pp):
        out = tr.TransportTrajectory.apply(pp, kk, L, Ls, Ug, c, dt, steps, every)
        d = (out - target)[inner]
        return 0.5 * (d * d).sum()

    kt = torch.from_numpy(k).cuda().requires_grad_(True)
    pt = torch.from_numpy(phi).cuda().requires_grad_(True)
    loss = loss_of(kt, pt)
    loss.backward()
    gk, gp = kt.grad.detach(), pt.grad.detach()
    eps = 1e-6
    with torch.no_grad():
        for j in range(4):
            v = torch.from_numpy(_periodic_random(L, rng)).cuda()
            for which, grad in (("k", gk), ("phi", gp)):
                if which == "k":
                    lp, lm = loss_of(kt + eps * v, pt), loss_of(kt - eps * v, pt)
                else:
                    lp, lm = loss_of(kt, pt + eps * v), loss_of(kt, pt - eps * v)
                fd = (lp - lm).item() / (2 * eps)
                an = (grad[inner] * v[inner]).sum().item()
                assert abs(fd - an) <= 1e-6 * max(1e-3, abs(an)), (which, j, fd, an)


def test_transport_errors():
    n = 8
    L, Ls, ops, k, phi, u, dt, c = _problem(n)
    Ug = [torch.from_numpy(x).cuda() for x in u]
    p = torch.from_numpy(phi).cuda()
    kk = torch.from_numpy(k).cuda()
    with pytest.raises(nh.NeboError, match="alias"):
        tr.stage(L, p, p, None, kk, Ug, c, dt, 1, face_layouts=Ls)
    with pytest.raises(nh.NeboError, match="stage must be"):
        tr.stage(L, torch.empty_like(p), p, p, kk, Ug, c, dt, 4, face_layouts=Ls)
    L2 = nh.Layout.build((n, n, n), 2)
    p2 = torch.zeros(L2.npts, dtype=torch.float64, device="cuda")
    with pytest.raises(nh.NeboError, match="memory window|ghost"):
        tr.stage(L, torch.empty_like(p), p, None, kk, [p2, Ug[1], Ug[2]], c, dt, 1,
                 face_layouts=[L2, Ls[1], Ls[2]])
    with pytest.raises(nh.NeboError, match="multiple of the checkpoint"):
        tr.forward(L, p, kk, Ug, c, dt, 7, 3, face_layouts=Ls)
