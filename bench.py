"""Benchmark: fused var-coefficient diffusion r <<= div(interp(k) * grad(phi)) on a
256^3 fp64 SVolField with one periodic ghost layer, on one MI355X.

One step = one evaluation of the fused assignment over the 256^3 interior with
inputs resident in HBM.  Prints ONE JSON line (the contract of this repository):
value = GCell-updates/s, plus the roofline of the dominant kernel (HIP events
around each launch on the launch stream) and a CPU baseline (the oracle's
restatement of the reference's serial and z-slab thread-parallel walks, timed on
this host).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--n 256] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spatialops-threadfix_amd", "python"))

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_PER_CELL = 24        # read phi, read k, write r (fp64)
METRIC = ("GCell-updates/s & %HBM peak, div(k∇φ) 256³ fp64; "
          "batched and adjoint vs plain")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-tune", action="store_true", help="keep the default tile (no autotune)")
    p.add_argument("--tile", default=None,
                   help="tx,ty,zchunk,variant of the diffusion kernel (implies --no-tune)")
    p.add_argument("--no-extras", action="store_true",
                   help="skip the batched, adjoint, transport, generic and heat legs")
    p.add_argument("--pmc-probe", action="store_true",
                   help="only launch 3 known-size copies and 3 kernels (for rocprofv3 --pmc runs)")
    p.add_argument("--pmc-probe-batched", action="store_true",
                   help="the same for the batched step: 3 copies of the patch blocks' size, 3 steps")
    p.add_argument("--batch-tile", default="0,2,16",
                   help="item tile of --pmc-probe-batched (x,y,z; x = 0: column items, y rows per lane)")
    return p.parse_args()


def _tile_name(t):
    """tx x ty x zchunk with the schedule: direct loads, LDS-staged + prefetch depth,
    or the column schedule (ty rows per lane)."""
    sched = {0: "", 1: "/staged-pf1", 2: "/staged-pf2", 3: "/column", 4: "/column-xreuse",
             5: "/stacked-ry4", 6: "/stacked-ry2", 7: "/stacked-ry4-pf3", 8: "/stacked-ry4-pf4",
             9: "/stacked-ry4-pf3-wb", 10: "/pair-ry4-pf3", 11: "/pair-ry2-pf3", 12: "/pair-ry4-pf2",
             13: "/pair-ry2-pf4", 14: "/pair-ry2-pf2", 15: "/pair-ry1-pf3", 16: "/pair-ry1-pf4"}[t[3]]
    return f"{t[0]}x{t[1]}x{t[2]}{sched}"


def cpu_baseline(L, kh, ph, coefs, budget_s):
    """The oracle (CPU restatement of the reference walk, built -O3 -DNDEBUG
    -ffp-contract=off as the reference's Release build) on this host, on the full
    256^3 workload: serial, and the z-slab thread-parallel split for P = 1, 2, 4,
    ... up to 16 threads (or the usable CPUs, if fewer).  Two restatements: the
    closed form (the per-cell arithmetic the reference's expression templates
    compile to, straight-line) and the generic tree walk (the same operations
    through the oracle's recursive evaluator of the expression tree: an
    interpreter, so slower than the compiled templates).  `value` is the fastest
    closed-form rate; `threads` holds both curves."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import nebo_hip as nh
    import oracle_lib as orc
    n3 = L.interior[0] * L.interior[1] * L.interior[2]
    r = nh.Field.from_numpy(L, np.zeros(L.npts), where="cpu")
    k = nh.Field(L, host=kh)
    phi = nh.Field(L, host=ph)
    n = L.interior
    ops = nh.basic_op_types(nh.build_stencils(n[0], n[1], n[2], 1.0, 1.0, 1.0))
    tree = (ops["DivX"](ops["InterpC2FX"](k) * ops["GradX"](phi))
            + ops["DivY"](ops["InterpC2FY"](k) * ops["GradY"](phi))
            + ops["DivZ"](ops["InterpC2FZ"](k) * ops["GradZ"](phi)))
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    top = max(1, min(16, usable))
    plist = [p for p in (1, 2, 4, 8, 16) if p <= top]
    if plist[-1] != top:
        plist.append(top)

    def rate(fn, min_s):
        reps, t0 = 0, time.perf_counter()
        while True:
            fn()
            reps += 1
            el = time.perf_counter() - t0
            if el > min_s or reps >= 50:
                return n3 * reps / el / 1e9

    per_point = budget_s / (2 * (len(plist) + 1))
    closed = {"serial": rate(lambda: orc.divkgrad(r, k, phi, coefs, 0), per_point)}
    for p in plist:
        closed[str(p)] = rate(lambda: orc.divkgrad(r, k, phi, coefs, 0, threads=p), per_point)
    ref = r.host.copy()
    walk = {"serial": rate(lambda: orc.assign(r, tree), 0.0)}
    same = bool(np.array_equal(r.host, ref))
    for p in plist:
        walk[str(p)] = rate(lambda: orc.assign(r, tree, threads=p), 0.0)
        same = same and bool(np.array_equal(r.host, ref))
    best_p = max(plist, key=lambda p: closed[str(p)])
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    rnd = lambda d: {key: round(v, 4) for key, v in d.items()}
    return {"value": round(closed[str(best_p)], 4), "unit": "GCell-updates/s", "cores": best_p, "kind": "port",
            "sample": (f"full {n[0]}^3 div(k grad phi) evaluations by the oracle's closed form, "
                       f"z-slab split over P threads (P = {', '.join(map(str, plist))}), "
                       f"about {per_point:.1f} s per point; fastest P = {best_p}; "
                       f"tree walk one evaluation per point; host {cpu_model}, "
                       f"{os.cpu_count()} CPUs, {usable} usable"),
            "serial_value": round(closed["serial"], 4),
            "threads": {"closed_form": rnd(closed), "tree_walk": rnd(walk)},
            "tree_walk_equals_closed_form": same,
            "nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": cpu_model,
            "oracle_build": "gcc -O3 -DNDEBUG -ffp-contract=off"}


def batched_dims(count=512):
    """Interior sizes of the batched workload: each axis drawn from {16,24,32,48,64}
    by a seeded splitmix64 stream."""
    import numpy as np
    from nebo_hip import inputs
    sizes = np.array([16, 24, 32, 48, 64])
    idx = (inputs.splitmix64(4, 3 * count) % np.uint64(5)).astype(np.int64).reshape(count, 3)
    return [tuple(int(v) for v in sizes[i]) for i in idx]


def batched_leg(steps=10, count=512):
    """512 patches with interior sizes drawn from {16,24,32,48,64} per axis (seeded),
    Laplacian + Euler per step: one persistent launch per step captured 10 at a
    time in a hipGraph, against one launch per patch per step."""
    import torch
    import nebo_hip as nh
    dims = batched_dims(count)
    cells = sum(d[0] * d[1] * d[2] for d in dims)
    db = nh.build_stencils(64, 64, 64, 1.0, 1.0, 1.0)
    coefs = nh.axis_coefs(db)
    s = 0.1 / 64.0 ** 2
    a, b = nh.alloc_patches(dims)
    for x in a:
        x.normal_()
    B = nh.Batch(a, b, dims, coefs, s)
    tile = B.autotune()
    stream = torch.cuda.Stream()
    ev = lambda: torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        B.graph_build(steps, stream)
        B.graph_launch(stream)
        e0, e1 = ev(), ev()
        e0.record(stream)
        for _ in range(3):
            B.graph_launch(stream)
        e1.record(stream)
        stream.synchronize()
        graph_ms = e0.elapsed_time(e1) / (3 * steps)
        e0, e1 = ev(), ev()
        e0.record(stream)
        for i in range(steps):
            B.step(i & 1, stream)
        e1.record(stream)
        stream.synchronize()
        eager_ms = e0.elapsed_time(e1) / steps
        B.step_per_patch(0, stream)
        e0, e1 = ev(), ev()
        e0.record(stream)
        for i in range(steps):
            B.step_per_patch(i & 1, stream)
        e1.record(stream)
        stream.synchronize()
        per_patch_ms = e0.elapsed_time(e1) / steps
    B.close()
    # a step reads every patch block whole (ghosts included) and writes its
    # interior; the steps alternate a -> b -> a, so each step reads what the one
    # before wrote.  The streaming rate of that pattern: device copies between two
    # buffers of the blocks' total size, alternating direction.
    ghosted = sum((d[0] + 2) * (d[1] + 2) * (d[2] + 2) for d in dims)
    min_bytes = 8 * ghosted + 8 * cells
    del a, b
    x = torch.randn(ghosted, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = ev(), ev()
    e0.record()
    for i in range(20):
        (y if i % 2 == 0 else x).copy_(x if i % 2 == 0 else y)
    e1.record()
    torch.cuda.synchronize()
    copy_GBps = 16 * ghosted / (e0.elapsed_time(e1) / 20 * 1e-3) / 1e9
    del x, y
    return {"patches": count, "cells_per_step": cells, "tile": list(tile),
            "graph_ms_per_step": round(graph_ms, 5), "persistent_ms_per_step": round(eager_ms, 5),
            "per_patch_launch_ms_per_step": round(per_patch_ms, 5),
            "speedup_vs_per_patch": round(per_patch_ms / graph_ms, 2),
            "gcells_per_s": round(cells / (graph_ms * 1e-3) / 1e9, 2),
            "hbm_frac": round(16 * cells / (graph_ms * 1e-3) / HBM_PEAK, 4),
            "min_traffic_bytes": min_bytes,
            "min_traffic_GBps": round(min_bytes / (graph_ms * 1e-3) / 1e9, 1),
            "alternating_copy_GBps": round(copy_GBps, 1),
            "frac_of_alternating_copy": round(min_bytes / (graph_ms * 1e-3) / 1e9 / copy_GBps, 4)}


def adjoint_leg(n=128, steps=100, every=10):
    """100 explicit-Euler steps of div(k grad phi) at 128^3 (periodic) with
    checkpoints every 10 steps, then the checkpoint/recompute reverse sweep."""
    import numpy as np
    import torch
    import nebo_hip as nh
    from nebo_hip import inputs
    from nebo_hip import adjoint as adj
    L = nh.Layout.build((n, n, n), 1)
    db = nh.build_stencils(n, n, n, 1.0, 1.0, 1.0)
    c = nh.axis_coefs(db)
    k = torch.from_numpy(inputs.coefficient_field(n, 5).reshape(-1)).cuda()
    phi0 = torch.from_numpy(inputs.fourier_field(n, 3).reshape(-1)).cuda()
    dt = 0.1 * (1.0 / n) ** 2 / 2.0
    phiT, ckpt = adj.forward(L, phi0, k, c, dt, steps, every)
    lamT = phiT.clone()
    adj.backward(L, lamT, k, c, dt, steps, every, ckpt)
    torch.cuda.synchronize()
    ev = lambda: torch.cuda.Event(enable_timing=True)
    e0, e1, e2 = ev(), ev(), ev()
    e0.record()
    phiT, ckpt = adj.forward(L, phi0, k, c, dt, steps, every)
    e1.record()
    adj.backward(L, lamT, k, c, dt, steps, every, ckpt)
    e2.record()
    torch.cuda.synchronize()
    fwd, bwd = e0.elapsed_time(e1), e1.elapsed_time(e2)
    return {"n": n, "steps": steps, "checkpoint_every": every, "forward_ms": round(fwd, 3),
            "backward_ms": round(bwd, 3), "backward_over_forward": round(bwd / fwd, 3)}


def _time_calls(fn, stream, reps, warm=3):
    """Mean ms per call over reps back-to-back calls, HIP events on stream."""
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def ceiling_leg(n=256, reps=20):
    """The streaming ceiling on blocks of the benchmark's size ((n+2)^3 fp64):
    r = a + s*b by nebo_stream_triad, two read streams and one write stream as
    div(k grad phi) has, without halos or arithmetic.  Each variant (access width,
    store policy) timed over `reps` back-to-back launches; GB/s counts 24 B per
    point."""
    import ctypes as C
    import torch
    import nebo_hip as nh
    npts = (n + 2) ** 3
    g = torch.Generator(device="cuda").manual_seed(5)
    a = torch.randn(npts, generator=g, dtype=torch.float64, device="cuda")
    b = torch.randn(npts, generator=g, dtype=torch.float64, device="cuda")
    r = torch.empty_like(a)
    lib = nh.lib()
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)
    names = {0: "8B-nt", 1: "16B-nt", 2: "16B-writeback", 3: "2x16B-nt", 4: "16B-nt-offset8"}
    rates, ok = {}, True
    for v in range(5):
        off = 1 if v == 4 else 0   # variant 4: every array 8 bytes past a 16-byte boundary
        m = npts - off

        def call():
            rc = lib.nebo_stream_triad(C.c_void_p(r.data_ptr() + 8 * off), C.c_void_p(a.data_ptr() + 8 * off),
                                       C.c_void_p(b.data_ptr() + 8 * off), 0.5, m, v, sp)
            if rc:
                raise RuntimeError(lib.nebo_last_error().decode())
        ms = _time_calls(call, stream, reps)
        rates[names[v]] = round(24 * m / (ms * 1e-3) / 1e9, 1)
        ok = ok and bool(torch.equal(r[off:], (a + 0.5 * b)[off:]))
    best = max(rates, key=rates.get)
    return {"workload": f"r = a + s*b over 3 blocks of {n + 2}^3 fp64", "bytes_per_launch": 24 * npts,
            "GBps": rates, "best": best, "best_GBps": rates[best],
            "best_frac_of_peak": round(rates[best] * 1e9 / HBM_PEAK, 4), "results_correct": ok}


def generic_leg(k, phi, ops, reps=20):
    """Trees that have no hand-written kernel, on the same 256^3 fields: each
    evaluated by the kernel compiled at run time for its shape (stencil trees on
    the column schedule, and again with it off: `flattened_us`) and by the
    stack-program interpreter (assignments must agree bit for bit), plus
    reductions.  GB/s counts algorithmic bytes: 8 B per distinct field read per
    cell and 8 B per written cell; the div(k grad phi) tree is forced through the
    generic evaluator to compare with the hand-written kernel."""
    import ctypes as C
    import torch
    import nebo_hip as nh
    L = phi.layout
    n3 = L.interior[0] * L.interior[1] * L.interior[2]
    out = nh.Field.empty(L, "gpu")
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)
    lib = nh.lib()
    div = sum((ops[f"Div{a}"](ops[f"InterpC2F{a}"](k) * ops[f"Grad{a}"](phi)) for a in "YZ"),
              ops["DivX"](ops["InterpC2FX"](k) * ops["GradX"](phi)))
    lap = sum((ops[f"Div{a}"](ops[f"Grad{a}"](phi)) for a in "YZ"), ops["DivX"](ops["GradX"](phi)))
    trees = {"cond(phi > 0, phi*phi)(-phi)": (nh.cond(phi > 0.0, phi * phi)(-phi), 16),
             "phi*k + sin(phi)": (phi * k + nh.sin(phi), 24),
             "div(k grad phi), generic": (div, 24),
             "laplacian, generic": (lap, 16)}
    res = {"assign": {}, "reduce": {}}
    for name, (t, bpc) in trees.items():
        s, d = nh.serialize(t, "gpu"), out.desc("gpu")

        def call():
            rc = lib.nebo_assign_generic(C.byref(d), s.ref, 1, sp)
            if rc:
                raise RuntimeError(lib.nebo_last_error().decode())
        e, outs = {}, []
        for mode in ("compiled", "flattened", "interpreter"):
            nh.set_fused_jit(mode != "interpreter")
            nh.fused_jit_column(mode == "compiled")
            ms = _time_calls(call, stream, reps)
            e[mode + "_us"] = round(ms * 1e3, 2)
            e[mode + "_GBps"] = round(bpc * n3 / (ms * 1e-3) / 1e9, 1)
            outs.append(out.gpu.clone())
        nh.set_fused_jit(True)
        nh.fused_jit_column()
        e["bit_identical"] = bool(torch.equal(outs[0], outs[2]) and torch.equal(outs[1], outs[2]))
        res["assign"][name] = e
    dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    reds = {"max(phi)": (nh.E(phi), nh._abi.RED_MAX, 1, 8),
            "norm_interior(phi)": (nh.E(phi), nh._abi.RED_NORM, 0, 8),
            "sum(phi*k)": (phi * k, nh._abi.RED_SUM, 1, 16)}
    for name, (t, op, ug, bpc) in reds.items():
        s = nh.serialize(t, "gpu")

        def call():
            rc = lib.nebo_reduce_async(s.ref, op, ug, C.c_void_p(dev.data_ptr()), sp)
            if rc:
                raise RuntimeError(lib.nebo_last_error().decode())
        e, vals = {}, []
        for mode in ("compiled", "interpreter"):
            nh.set_fused_jit(mode == "compiled")
            ms = _time_calls(call, stream, reps)
            e[mode + "_us"] = round(ms * 1e3, 2)
            e[mode + "_GBps"] = round(bpc * n3 / (ms * 1e-3) / 1e9, 1)
            vals.append(float(dev.item()))
        nh.set_fused_jit(True)
        e["values"] = vals
        res["reduce"][name] = e
    return res


def heat_leg(steps=100, cpu_steps=5, per_graph=10):
    """The heat equation of examples/heat_equation.cpp at 64^3: each step
    phi' <<= phi + s*lap(phi) (hand-written Laplacian + Euler kernel), the
    zero-Dirichlet ghosts reset by phi' <<= cond(ghost mask, 0)(phi') (a
    compiled kernel) and field_max, field_sum of phi' into device memory (one
    kernel each, no host synchronisation inside the loop).  Timed with HIP events
    around `steps` eager steps, and around the same steps replayed from a hipGraph
    of `per_graph` captured steps (both runs must end in the same bits); the
    oracle runs `cpu_steps` of the same steps on the host."""
    import numpy as np
    import torch
    import nebo_hip as nh
    from nebo_hip import workloads as W
    n = 64
    H = W.HeatStepper(n, slots=steps)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for i in range(2):   # warm: compiles the cond and reduction kernels, creates this stream's buffer
            H.step(i, stream)
        H.reset()
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for i in range(steps):
            H.step(i, stream)
        e1.record(stream)
        stream.synchronize()
        gpu_ms = e0.elapsed_time(e1) / steps
        eager_phi = H.fields[steps % 2].gpu.clone()
        eager_red = H.red.clone()
        graph = torch.cuda.CUDAGraph()
        H.reset()
        stream.synchronize()
    with torch.cuda.graph(graph, stream=stream):
        for i in range(per_graph):
            H.step(i, stream)
    with torch.cuda.stream(stream):
        H.reset()
        H.red.zero_()
        stream.synchronize()
        g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        g0.record(stream)
        for _ in range(steps // per_graph):
            graph.replay()
        g1.record(stream)
        stream.synchronize()
        graph_ms = g0.elapsed_time(g1) / (steps // per_graph * per_graph)
    same = bool(torch.equal(H.fields[steps % 2].gpu, eager_phi)
                and torch.equal(H.red[:2 * per_graph], eager_red[2 * (steps - per_graph):2 * steps]))
    maxes = eager_red[0::2].cpu().numpy()
    # the same steps on the host with the oracle (the reference's CPU backend)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    L, _, s, mask = W.heat_problem(n)
    h = [nh.Field.from_numpy(L, H.phi0, where="cpu"), nh.Field.from_numpy(L, H.phi0, where="cpu")]
    c = nh.axis_coefs(nh.build_stencils(n, n, n, 1.0, 1.0, 1.0))
    t0 = time.perf_counter()
    for i in range(cpu_steps):
        src, dst = h[i & 1], h[1 - (i & 1)]
        orc.lapl_euler(dst, src, c, s)
        orc.assign(dst, nh.cond(mask, 0.0)(dst))
        orc.reduce(dst, nh._abi.RED_MAX)
        orc.reduce(dst, nh._abi.RED_SUM)
    cpu_ms = (time.perf_counter() - t0) * 1e3 / cpu_steps
    return {"n": n, "steps": steps, "gpu_ms_per_step": round(gpu_ms, 4),
            "cpp_front_end": cpp_heat_timing(n),
            "graph_ms_per_step": round(graph_ms, 4), "steps_per_graph": per_graph,
            "graph_equals_eager": same,
            "kernels_per_step": "Laplacian+Euler, cond mask (one kernel each), max, sum (block partials, then a fold: two kernels each)",
            "max_first_last": [float(maxes[0]), float(maxes[-1])],
            "max_decreasing": bool(np.all(np.diff(maxes) <= 0)),
            "oracle_cpu_ms_per_step_serial": round(cpu_ms, 3), "cpu_steps_timed": cpu_steps}


def cpp_heat_timing(n, steps=200):
    """The time loop of examples/heat_equation.cpp (three assignments per step,
    the z terms added) at n^3 through the C++ front end: cpp/bin/heat_timing, run
    as a child process; its JSON line, or a note when it is not built or fails."""
    import subprocess
    exe = os.path.join(ROOT, "spatialops-threadfix_amd", "cpp", "bin", "heat_timing")
    if not os.path.exists(exe):
        return {"skipped": "cpp/bin/heat_timing not built"}
    r = subprocess.run([exe, str(n), str(steps)], capture_output=True, text=True, timeout=120)
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"failed": r.returncode, "stderr": r.stderr[-300:]}
    return json.loads(lines[-1])


def particles_leg(n=128, count=4_000_000, reps=5):
    """Particle <-> cell interpolation (ParticleToCell / CellToParticle of the
    reference's particle operators) on a 128^3 cell field, 4M particles with
    seeded uniform positions and cube edges of 0.2-1.5 cells: each C2P gathers
    the volume-weighted average over the cells a particle overlaps, each P2C
    spreads a particle quantity over them (count, scan, fill, stable radix sort
    of (cell, value) pairs, ordered run sums: bit-identical to the serial loop)."""
    import torch
    import nebo_hip as nh
    from nebo_hip import particles as P
    L = nh.Layout.build((n, n, n), 1)
    h = 1.0 / n
    geom = (h, 0.5 * h, h, 0.5 * h, h, 0.5 * h)
    g = torch.Generator(device="cuda").manual_seed(11)
    size = (0.2 + 1.3 * torch.rand(count, generator=g, device="cuda", dtype=torch.float64)) * h
    pos = [size / 2 + (1.0 - size) * torch.rand(count, generator=g, device="cuda", dtype=torch.float64)
           for _ in range(3)]
    vals = torch.randn(count, generator=g, device="cuda", dtype=torch.float64)
    cell = nh.Field.empty(L, "gpu")
    cell.gpu.normal_(generator=g)
    dest = torch.empty(count, dtype=torch.float64, device="cuda")
    c2p, p2c = P.CellToParticle(*geom), P.ParticleToCell(*geom)
    for op in (c2p, p2c):
        op.set_coordinate_information(pos[0], pos[1], pos[2], size)
    out = nh.Field.empty(L, "gpu")

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    c2p_ms = timed(lambda: c2p.apply_to_field(cell, dest))

    p2c_ms = timed(lambda: p2c.apply_to_field(vals, out))   # zeroes out first, as the reference does
    # conservation: every particle's value lands in full in the cells it overlaps
    total = float(out.gpu.sum().item())
    ref = float(vals.sum().item())
    return {"cells": n ** 3, "particles": count, "c2p_ms": round(c2p_ms, 4), "p2c_ms": round(p2c_ms, 4),
            "c2p_mparticles_per_s": round(count / (c2p_ms * 1e-3) / 1e6, 1),
            "p2c_mparticles_per_s": round(count / (p2c_ms * 1e-3) / 1e6, 1),
            "p2c_sum_rel_err": abs(total - ref) / max(1e-300, float(vals.abs().sum().item()))}


def transport_leg(n=512, steps=20, every=5):
    """Scalar transport at 512^3: upwinded convective flux (cond on the sign of the
    face velocity) plus div(k grad phi), SSP-RK3, 20 steps.  Timed two ways: each
    stage as one fused assignment on the generic evaluator, and on the hand-written
    stage kernel (the two must agree bit for bit); then the adjoint of the 20 steps
    with respect to k and phi0 with checkpoints every 5 steps."""
    import torch
    import nebo_hip as nh
    from nebo_hip import workloads as W
    from nebo_hip import transport as tr
    L, Ls, ops, k, phi, u, dt = W.transport_problem_gpu(n)
    c = nh.axis_coefs(nh.build_stencils(n, n, n, 1.0, 1.0, 1.0))
    K = nh.Field(L, gpu=k)
    U = [nh.Field(l, gpu=x) for l, x in zip(Ls, u)]
    P = [nh.Field(L, gpu=phi.clone()), nh.Field(L, gpu=phi.clone()), nh.Field(L, gpu=phi.clone())]
    lo, hi = W.upwind_ops()
    s1, s2, s3 = W.rk3_stage_exprs(ops, P[0], K, U, dt, lo, hi)
    e1, e2, e3 = s1(), s2(P[1]), s3(P[2])
    ev = lambda: torch.cuda.Event(enable_timing=True)
    abs0 = nh.field_sum_interior(nh.abs_(P[0]))
    sum0 = nh.field_sum_interior(P[0])

    def generic_step():
        nh.assign(P[1], e1)
        nh.fill_ghosts_periodic(P[1])
        nh.assign(P[2], e2)
        nh.fill_ghosts_periodic(P[2])
        nh.assign(P[0], e3)
        nh.fill_ghosts_periodic(P[0])

    H = [phi.clone(), torch.empty_like(phi), torch.empty_like(phi), torch.empty_like(phi)]

    def hand_step(cur):
        nxt = 3 - cur if cur in (0, 3) else 0
        tr.stage(L, H[1], H[cur], None, k, u, c, dt, 1, face_layouts=Ls)
        tr.stage(L, H[2], H[cur], H[1], k, u, c, dt, 2, face_layouts=Ls)
        tr.stage(L, H[nxt], H[cur], H[2], k, u, c, dt, 3, face_layouts=Ls)
        return nxt

    generic_step()
    cur = hand_step(0)
    torch.cuda.synchronize()
    a0, a1, b0, b1 = ev(), ev(), ev(), ev()
    a0.record()
    for _ in range(steps - 1):
        generic_step()
    a1.record()
    b0.record()
    for _ in range(steps - 1):
        cur = hand_step(cur)
    b1.record()
    torch.cuda.synchronize()
    gen_ms = a0.elapsed_time(a1) / (steps - 1)
    hand_ms = b0.elapsed_time(b1) / (steps - 1)
    same = bool(torch.equal(P[0].gpu, H[cur]))
    sum1 = nh.field_sum_interior(P[0])
    # adjoint: forward with checkpoints, then the reverse sweep
    phiT, ckpt = tr.forward(L, phi, k, u, c, dt, steps, every, face_layouts=Ls)
    tr.backward(L, phiT.clone(), k, u, c, dt, steps, every, ckpt, face_layouts=Ls)
    torch.cuda.synchronize()
    f0, f1, f2 = ev(), ev(), ev()
    f0.record()
    phiT, ckpt = tr.forward(L, phi, k, u, c, dt, steps, every, face_layouts=Ls)
    f1.record()
    lam0, gk = tr.backward(L, phiT.clone(), k, u, c, dt, steps, every, ckpt, face_layouts=Ls)
    f2.record()
    torch.cuda.synchronize()
    fwd_ms, bwd_ms = f0.elapsed_time(f1), f1.elapsed_time(f2)
    cells = n ** 3
    return {"n": n, "steps": steps,
            "generic_ms_per_rk3_step": round(gen_ms, 3),
            "kernel_ms_per_rk3_step": round(hand_ms, 3),
            "kernel_gcells_per_s_per_stage": round(3 * cells / (hand_ms * 1e-3) / 1e9, 2),
            "kernel_stage_hbm_frac": round(3 * 56 * cells / (hand_ms * 1e-3) / HBM_PEAK, 4),
            "generic_stage_hbm_frac": round(3 * 56 * cells / (gen_ms * 1e-3) / HBM_PEAK, 4),
            "kernel_equals_generic": same,
            "sum_drift_over_sum_abs": abs(sum1 - sum0) / abs0,
            "program_insns": nh.lower_stats(e3, where="gpu")[0],
            "adjoint": {"checkpoint_every": every, "forward_ms": round(fwd_ms, 2),
                        "backward_ms": round(bwd_ms, 2), "backward_over_forward": round(bwd_ms / fwd_ms, 3)}}


def pmc_probe(n, tile=None):
    """Three copies of one field (known bytes) and three fused kernel launches,
    nothing else: the counter runs divide the kernel's FETCH/WRITE by the copy's."""
    import torch
    import nebo_hip as nh
    from nebo_hip import inputs
    L = nh.Layout.build((n, n, n), 1)
    k = nh.Field.from_numpy(L, inputs.coefficient_field(n, 2).reshape(-1), where="gpu")
    phi = nh.Field.from_numpy(L, inputs.fourier_field(n, 3).reshape(-1), where="gpu")
    r = nh.Field.empty(L, "gpu")
    dst = torch.empty_like(phi.gpu)
    db = nh.build_stencils(n, n, n, 1.0, 1.0, 1.0)
    c = nh.axis_coefs(db)
    if tile:
        nh.stencil_set_tile(0, *tile)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(phi.gpu)
    for _ in range(3):
        nh.divkgrad(r, k, phi, c, 0)
    # the streaming ceiling's kernels on the same blocks (every triad variant twice)
    import ctypes as C
    sp = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for v in (0, 1, 2, 3, 4, 0, 1, 2, 3, 4):
        off = 8 if v == 4 else 0   # variant 4: every array 8 bytes past a 16-byte boundary
        nh._abi.check(nh.lib().nebo_stream_triad(C.c_void_p(dst.data_ptr() + off),
                                                 C.c_void_p(phi.gpu.data_ptr() + off),
                                                 C.c_void_p(k.gpu.data_ptr() + off), 0.5, L.npts - off // 8, v, sp))
    torch.cuda.synchronize()
    print(json.dumps({"pmc_probe": True, "copy_bytes_read": L.npts * 8, "copy_bytes_written": L.npts * 8,
                      "kernel_bytes_algorithmic": BYTES_PER_CELL * n ** 3,
                      "tile": _tile_name(nh.stencil_select_tile(0, L.interior))}))


def pmc_probe_batched(tile, count=512):
    """Three copies of a buffer the size of the patch blocks (known bytes) and three
    batched steps (alternating a -> b -> a) with the given item tile, nothing else."""
    import torch
    import nebo_hip as nh
    dims = batched_dims(count)
    ghosted = sum((d[0] + 2) * (d[1] + 2) * (d[2] + 2) for d in dims)
    cells = sum(d[0] * d[1] * d[2] for d in dims)
    db = nh.build_stencils(64, 64, 64, 1.0, 1.0, 1.0)
    a, b = nh.alloc_patches(dims)
    for x in a:
        x.normal_()
    B = nh.Batch(a, b, dims, nh.axis_coefs(db), 0.1 / 64.0 ** 2, tile=tile)
    x = torch.randn(ghosted, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    for i in range(3):
        B.step(i & 1)
    torch.cuda.synchronize()
    info = B.info()
    B.close()
    print(json.dumps({"pmc_probe_batched": True, "copy_bytes_read": 8 * ghosted, "copy_bytes_written": 8 * ghosted,
                      "step_bytes_read_min": 8 * ghosted, "step_bytes_written_min": 8 * cells,
                      "tile": list(info[0]), "items": info[1]}))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    tile_arg = tuple(int(v) for v in a.tile.split(",")) if a.tile else None
    if a.pmc_probe:
        pmc_probe(a.n, tile_arg)
        return
    if a.pmc_probe_batched:
        pmc_probe_batched(tuple(int(v) for v in a.batch_tile.split(",")))
        return
    import numpy as np
    import torch
    import nebo_hip as nh
    from nebo_hip import inputs

    n = a.n
    L = nh.Layout.build((n, n, n), 1)
    kh = inputs.coefficient_field(n, 2).reshape(-1)
    ph = inputs.fourier_field(n, 3).reshape(-1)
    k = nh.Field.from_numpy(L, kh, where="gpu")
    phi = nh.Field.from_numpy(L, ph, where="gpu")
    r = nh.Field.empty(L, "gpu")
    r.gpu.zero_()
    db = nh.build_stencils(n, n, n, 1.0, 1.0, 1.0)
    ops = nh.basic_op_types(db)
    coefs = nh.axis_coefs(db)
    tree = (ops["DivX"](ops["InterpC2FX"](k) * ops["GradX"](phi))
            + ops["DivY"](ops["InterpC2FY"](k) * ops["GradY"](phi))
            + ops["DivZ"](ops["InterpC2FZ"](k) * ops["GradZ"](phi)))
    assert nh.assign_plan(r, tree) == nh._abi.KERNEL_DIVKGRAD
    # the serialised descriptor is reused across steps, as a caller's plan would be
    ser = nh.serialize(tree, "gpu")
    import ctypes as C
    rd = r.desc("gpu")
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)
    lib = nh.lib()

    def step():
        rc = lib.nebo_assign(C.byref(rd), ser.ref, 1, sp)
        if rc:
            nh._abi.check(rc)

    tune = None
    if tile_arg:
        nh.stencil_set_tile(0, *tile_arg)
    elif not a.no_tune:
        best, times = nh.stencil_autotune(r, k, phi, coefs)
        tune = {"tile": _tile_name(best),
                "ms_per_tile": {_tile_name(t): round(v, 5) for t, v in times.items()}}
    tile = nh.stencil_select_tile(0, L.interior)   # the tile the launches use (pinned, tuned or by size)
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    # whole-job timing: K back-to-back steps between two synchronisations, HIP
    # events on the launch stream around them.  Each step is one kernel launch,
    # so the mean step is the kernel's mean duration in steady state (an event
    # between two launches would add a pipeline bubble of several us to each).
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    g0.record(stream)
    for _ in range(a.steps):
        step()
    g1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = g0.elapsed_time(g1) / a.steps
    avg_kernel_ms = ms_per_step
    cells = n ** 3
    gcells = cells / (ms_per_step * 1e-3) / 1e9
    achieved = BYTES_PER_CELL * cells / (avg_kernel_ms * 1e-3)

    # HBM bytes per launch measured with rocprofv3 counters for this tile (null if
    # this tile has not been measured)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"divkgrad_{n}_traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("hbm_bytes_per_launch", {}).get(_tile_name(tile))
        except (OSError, ValueError, AttributeError):
            traffic = None

    ceiling = ceiling_leg(n)
    extras = {}
    if not a.no_extras:
        extras["batched"] = batched_leg()
        extras["adjoint"] = adjoint_leg()
        extras["transport"] = transport_leg()
        extras["heat"] = heat_leg()
        extras["particles"] = particles_leg()
        extras["generic"] = generic_leg(k, phi, ops)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(L, kh, ph, coefs, a.cpu_seconds)

    line = {
        "metric": METRIC, "value": round(gcells, 3), "unit": "GCell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded k ~ U[0.5,2), 8-mode Fourier phi, periodic ghosts)",
        "config": {"workload": f"div(k grad phi) {n}^3 fp64, 1 ghost layer, fused <<=",
                   "n": n, "cells": cells, "kernel": "divkgrad (hand-written HIP, gfx950)",
                   "tile": _tile_name(tile), "tile_args": ",".join(map(str, tile)), "autotune": tune},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "bytes_per_launch_algorithmic": BYTES_PER_CELL * cells,
                     "avg_kernel_ms": round(avg_kernel_ms, 5),
                     "hbm_GBps_measured_traffic": (round(traffic / (avg_kernel_ms * 1e-3) / 1e9, 1)
                                                   if traffic else None),
                     "stream_ceiling_GBps": ceiling["best_GBps"],
                     "frac_of_stream_ceiling": round(achieved / 1e9 / ceiling["best_GBps"], 4)},
        "stream_ceiling": ceiling,
        "cpu_baseline": cpu,
        **extras,
        "wall_s_timed_region": round(wall, 4),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
