"""The benchmark workloads end to end on the GPU, against the oracle:
  * heat equation (64^3, zero-Dirichlet ghosts set by a mask, max/sum each step)
  * scalar transport (upwind via cond on the face velocity + diffusion, SSP-RK3)
Fields are compared bit for bit after every step; max exactly; sums within the
serial sum's own error bound."""
import math

import numpy as np
import pytest
import torch

import nebo_hip as nh
from nebo_hip import workloads as W, _abi
import oracle_lib as orc

pytestmark = pytest.mark.gpu


def test_heat_equation_trajectory():
    n, steps = 64, 100
    L, ops, s, mask = W.heat_problem(n)
    mask.add_consumer(0)
    phi0 = W.heat_initial(n)
    g = [nh.Field.from_numpy(L, phi0, where="gpu"), nh.Field.from_numpy(L, phi0, where="gpu")]
    h = [nh.Field.from_numpy(L, phi0, where="cpu"), nh.Field.from_numpy(L, phi0, where="cpu")]
    db = nh.build_stencils(n, n, n, 1.0, 1.0, 1.0)
    c = nh.axis_coefs(db)
    for step in range(steps):
        src, dst = step % 2, (step + 1) % 2
        e1, cond_of = W.heat_step_exprs(ops, g[src], s, mask)
        assert step or nh.assign_plan(g[dst], e1) == _abi.KERNEL_LAPL_EULER
        nh.assign(g[dst], e1)
        nh.assign(g[dst], cond_of(g[dst]))
        mx, sm = nh.field_max(g[dst]), nh.field_sum(g[dst])
        orc.lapl_euler(h[dst], h[src], c, s)
        orc.assign(h[dst], nh.cond(mask, 0.0)(h[dst]))
        assert mx == orc.reduce(h[dst], _abi.RED_MAX)
        a = h[dst].host
        assert abs(sm - math.fsum(a)) <= 2.0 ** -52 * abs(math.fsum(a)) + 1e-300
        if step % 25 == 24 or step == steps - 1:
            torch.cuda.synchronize()
            assert np.array_equal(g[dst].gpu.cpu().numpy(), h[dst].host), f"step {step}"
    # the ghosts are zero and the heat spreads: max decreases monotonically from the start
    assert nh.field_max(g[steps % 2]) < phi0.max()


def test_transport_rk3_bit_exact():
    n, steps = 20, 2
    L, Ls, ops, k, phi, u, dt = W.transport_problem(n)
    lo, hi = W.upwind_ops()

    def fields(where):
        K = nh.Field.from_numpy(L, k, where=where)
        U = [nh.Field.from_numpy(l, x, where=where) for l, x in zip(Ls, u)]
        P = [nh.Field.from_numpy(L, phi, where=where) for _ in range(3)]
        return K, U, P

    Kg, Ug, Pg = fields("gpu")
    Kh, Uh, Ph = fields("cpu")
    for step in range(steps):
        for (K, U, P, run) in ((Kg, Ug, Pg, nh.assign), (Kh, Uh, Ph, orc.assign)):
            s1, s2, s3 = W.rk3_stage_exprs(ops, P[0], K, U, dt, lo, hi)
            run(P[1], s1())
            (nh.fill_ghosts_periodic if run is nh.assign else orc.fill_periodic)(P[1])
            run(P[2], s2(P[1]))
            (nh.fill_ghosts_periodic if run is nh.assign else orc.fill_periodic)(P[2])
            run(P[0], s3(P[2]))
            (nh.fill_ghosts_periodic if run is nh.assign else orc.fill_periodic)(P[0])
        torch.cuda.synchronize()
        assert np.array_equal(Pg[0].gpu.cpu().numpy(), Ph[0].host), f"step {step}"
    # conservation: the periodic flux-form update keeps the interior sum
    a0 = phi.reshape(L.glob[::-1])[1:-1, 1:-1, 1:-1]
    a1 = Ph[0].host.reshape(L.glob[::-1])[1:-1, 1:-1, 1:-1]
    assert abs(math.fsum(a1.ravel()) - math.fsum(a0.ravel())) < 1e-10 * np.abs(a0).sum()


def test_transport_velocity_is_divergence_free():
    n = 16
    L, Ls, ops, k, phi, u, dt = W.transport_problem(n)
    U = [nh.Field.from_numpy(l, x, where="gpu") for l, x in zip(Ls, u)]
    d = nh.Field.empty(L, "gpu")
    nh.assign(d, ops["DivX"](U[0]) + ops["DivY"](U[1]) + ops["DivZ"](U[2]))
    torch.cuda.synchronize()
    inner = d.gpu.cpu().numpy().reshape(L.glob[::-1])[1:-1, 1:-1, 1:-1]
    assert np.abs(inner).max() < 1e-9 * n


def test_heat_step_graph_replays_match_eager():
    """Ten heat steps (Laplacian + Euler, cond mask, max and sum into device
    memory) captured in a hipGraph once the stream has run a step, replayed three
    times with the same steps run eagerly on other fields and host
    synchronisations between the replays: after every replay the field and the
    max/sum values equal the eager run's bit for bit."""
    per = 10
    H = W.HeatStepper(64, slots=per)
    Eg = W.HeatStepper(64, slots=per)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        H.step(0, stream)
        H.step(1, stream)
        H.reset()
        stream.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        for i in range(per):
            H.step(i, stream)
    with torch.cuda.stream(stream):
        H.reset()
        Eg.reset()
        stream.synchronize()
        for rep in range(3):
            print(f"heat graph replay {rep}", flush=True)
            graph.replay()
            for i in range(per):
                Eg.step(i, stream)
            stream.synchronize()
            assert torch.equal(H.fields[0].gpu, Eg.fields[0].gpu), rep
            assert torch.equal(H.red, Eg.red), rep
    mx = H.red[0::2].cpu().numpy()
    assert np.all(np.diff(mx) <= 0)
