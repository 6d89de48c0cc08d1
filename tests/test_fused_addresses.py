"""Address audit of the column-schedule kernels the run-time compiler generates.

Those kernels read each field through a buffer resource per plane (its range
ends at the field's last valid element) at per-lane byte offsets, with the
values past a wave's left and right edges loaded once per family.  A resource
whose base or offset reaches below a field's allocation reads memory the field
does not own, which on the GPU is a memory access fault.

Here each generated kernel is compiled for the host (tests/jit_host_shim.h
stands in for the device builtins) and run lane by lane over every block of the
launch the library would make (nebo_fused_column_launch gives the source, the
argument struct and the grid), with host fields in place of device ones.  Every
buffer load inside its resource's range must address one of the fields'
allocations, every store must land in the assigned field, and the stores must
cover the assignment's region exactly once per point.  The trees are random
stencil chains over operators that read both sides of a point, one side only,
and two points away, on fields with one and two ghost layers, asymmetric and
partly invalid ghosts, boxes narrower and wider than a wave, and several block
shapes."""
import ctypes as C
import os
import random
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import nebo_hip as nh
from nebo_hip import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
SHIM = os.path.join(HERE, "jit_host_shim.h")

DRIVER = r'''
static_assert(sizeof(JitArgs) == 1280, "argument struct size");
static const char* g_out_lo = nullptr;
static long g_out_n = 0;
static long* g_hits = nullptr;
static long g_bad_stores = 0;

extern "C" long audit_run(const void* args, int grid, int block, const void* const* lo, const void* const* hi,
                          int nr, const void* out_lo, long out_n, long* hits, long* counts)
{
  g_nranges = nr;
  for (int i = 0; i < nr; ++i) g_ranges[i] = ShimRange{(const char*)lo[i], (const char*)hi[i]};
  g_out_lo = (const char*)out_lo;
  g_out_n = out_n;
  g_hits = hits;
  g_bad = g_loads = g_stores = g_bad_stores = 0;
  g_first_bad = nullptr;
  JitArgs A;
  std::memcpy(&A, args, sizeof A);
  for (int b = 0; b < grid; ++b)
    for (int t = 0; t < block; ++t) {
      blockIdx.x = b;
      threadIdx.x = t;
      nebo_fused(A);
    }
  counts[0] = g_loads;
  counts[1] = g_stores;
  counts[2] = (long)(g_first_bad ? g_first_bad - (const char*)0 : 0);
  counts[3] = g_bad_stores;
  return g_bad;
}

static void shim_store_hit(const void* p, size_t bytes)
{
  const char* c = (const char*)p;
  if (c < g_out_lo || c + bytes > g_out_lo + 8 * g_out_n || (c - g_out_lo) % 8) {
    ++g_bad_stores;
    return;
  }
  for (size_t b = 0; b < bytes; b += 8) ++g_hits[(c - g_out_lo + b) / 8];
}
'''

GXX = shutil.which("g++")
pytestmark = pytest.mark.skipif(GXX is None, reason="g++ not installed")


def column_launch(lhs, rhs, use_ghost):
    """(source, argument bytes, grid, block) of the column-schedule launch, or None."""
    s = nh.serialize(nh.E(rhs), "host")
    d = lhs.desc("host")
    need = C.c_int32()
    args = C.create_string_buffer(1280)
    grid, block = C.c_int32(), C.c_int32()
    rc = _abi.lib().nebo_fused_column_launch(C.byref(d), s.ref, int(use_ghost), None, 0, C.byref(need), args,
                                             1280, C.byref(grid), C.byref(block))
    if rc == _abi.NEBO_ERR_UNSUPPORTED:
        return None
    assert rc == 0, _abi.lib().nebo_last_error()
    buf = C.create_string_buffer(need.value)
    _abi.call("nebo_fused_column_launch", C.byref(d), s.ref, int(use_ghost), buf, need.value, C.byref(need), args,
              1280, C.byref(grid), C.byref(block))
    return buf.value.decode(), args.raw, grid.value, block.value


def build(sources, tmp):
    """Compile each generated kernel with the shim and the driver, in parallel."""
    procs, libs = [], []
    for i, src in enumerate(sources):
        cpp = os.path.join(tmp, f"k{i}.cpp")
        so = os.path.join(tmp, f"k{i}.so")
        with open(cpp, "w") as fh:
            fh.write('#include "jit_host_shim.h"\n')
            fh.write(src)
            fh.write(DRIVER)
        procs.append(subprocess.Popen([GXX, "-std=c++17", "-O1", "-w", "-shared", "-fPIC", "-I", HERE, cpp, "-o",
                                       so], stderr=subprocess.PIPE, text=True))
        libs.append(so)
        if len(procs) >= 8:
            for p in procs:
                assert p.wait() == 0, p.stderr.read()
            procs = []
    for p in procs:
        assert p.wait() == 0, p.stderr.read()
    return libs


def audit(so, fields, lhs, launch, lim):
    src, args, grid, block = launch
    lib = C.CDLL(so)
    lib.audit_run.restype = C.c_long
    lo = (C.c_void_p * len(fields))(*[f.host.ctypes.data for f in fields])
    hi = (C.c_void_p * len(fields))(*[f.host.ctypes.data + f.host.nbytes for f in fields])
    hits = np.zeros(lhs.host.size, dtype=np.int64)
    counts = (C.c_long * 4)()
    bad = lib.audit_run(args, grid, block, lo, hi, len(fields), C.c_void_p(lhs.host.ctypes.data), lhs.host.size,
                        hits.ctypes.data_as(C.c_void_p), counts)
    where = ""
    if bad:
        addr = counts[2]
        for f in fields:
            base = f.host.ctypes.data
            where += f" [{f.name}: {(addr - base) / 8:+.0f} elements from its start]"
    assert bad == 0, f"{bad} loads outside the fields{where}"
    assert counts[3] == 0, f"{counts[3]} stores outside the assigned field"
    g = lhs.layout.glob
    want = np.zeros((g[2], g[1], g[0]), dtype=np.int64)
    (l0, l1, l2), (h0, h1, h2) = lim
    # limits count from the field's valid region: it starts offset + valid_minus
    # into the glob block
    gm = [lhs.layout.offset[d] + lhs.layout.valid_minus[d] for d in range(3)]
    want[l2 + gm[2]:h2 + gm[2], l1 + gm[1]:h1 + gm[1], l0 + gm[0]:h0 + gm[0]] = 1
    assert np.array_equal(hits.reshape(want.shape), want), "stores do not cover the region exactly once"
    assert counts[0] > 0


def host_field(L, name, rng):
    f = nh.Field.empty(L, "cpu", name=name)
    f.host[:] = rng.standard_normal(L.npts)
    return f


def stencil_tree(rng, leaves, stencils, depth):
    if depth == 0 or rng.random() < 0.25:
        return nh.E(rng.uniform(-2, 2)) if rng.random() < 0.1 else nh.E(rng.choice(leaves))
    sub = lambda: stencil_tree(rng, leaves, stencils, depth - 1)
    k = rng.random()
    if k < 0.5:
        return rng.choice(stencils)(sub())
    if k < 0.85:
        return nh.Expr(rng.choice(["ADD", "SUB", "MUL", "MAX"]), (sub(), sub()))
    return nh.Expr(rng.choice(["NEG", "ABS", "SQRT"]), (sub(),))


def cases():
    """(layout, use_ghost, column shape, seed) combinations."""
    out = []
    # mode 1: two columns per lane where the register ring allows, 2: one column
    specs = [((13, 9, 11), 2, (1, 1, 0)), ((13, 9, 11), 1, (1, 2, 1)), ((70, 6, 5), 2, (1, 1, 3)),
             ((130, 3, 17), 1, (1, 4, 0)), ((64, 4, 8), ((1, 2, 1), (2, 1, 2)), (1, 2, 5)),
             ((5, 13, 7), 2, (1, 1, 2)), ((66, 10, 4), "valid", (1, 1, 0)),
             ((13, 9, 11), 2, (2, 1, 0)), ((70, 6, 5), 2, (2, 1, 3)), ((130, 3, 17), 1, (2, 4, 0)),
             ((66, 10, 4), "valid", (2, 1, 0)), ((129, 5, 6), 1, (1, 0, 0)), ((127, 4, 9), 2, (1, 1, 4)),
             ((200, 11, 7), 2, (1, 0, 3))]
    for i, (n, ghost, shape) in enumerate(specs):
        for use_ghost in (True, False):
            out.append((n, ghost, use_ghost, shape, 100 + 2 * i + int(use_ghost)))
    return out


def test_generated_column_kernels_stay_in_bounds():
    stencil_names = [("Interpolant", "SVol", "XSurfY"), ("GradientY", "SVol", "SVol"),
                     ("Gradient", "SVol", "SSurfX"), ("Divergence", "SSurfX", "SVol"),
                     ("Divergence", "SSurfZ", "SVol"), ("Interpolant", "SVol", "SSurfY"),
                     ("GradientX", "SVol", "SVol"), ("Interpolant", "SSurfZ", "SVol"), ("Filter", "SVol", "SVol")]
    jobs = []
    try:
        for n, ghost, use_ghost, shape, seed in cases():
            nh.fused_jit_column(*shape)
            rng = np.random.default_rng(seed)
            if ghost == "valid":
                L = nh.Layout.build(n, 2).with_valid((1, 2, 0), (2, 1, 1))
            else:
                L = nh.Layout.build(n, ghost)
            f, g, h = host_field(L, "f", rng), host_field(L, "g", rng), host_field(L, "h", rng)
            out = nh.Field.empty(L, "cpu", name="out")
            db = nh.build_stencils(*n, 1.0, 1.0, 1.0)
            stencils = [db[k] for k in stencil_names] + [nh.Stencil([(1, 0, 0)], [2.0], kind=_abi.STENCIL_EDGELESS),
                                                         nh.Stencil([(-2, 1, 0), (0, -1, 2)], [0.5, 1.5])]
            r = random.Random(seed)
            found = 0
            for _ in range(40):
                t = stencil_tree(r, [f, g, h], stencils, 4)
                try:
                    nh.lower_stats(t, "host")
                    lim = nh.assign_limits(out, t, use_ghost, where="host")
                except nh.NeboError:
                    continue
                if lim is None:
                    continue
                launch = column_launch(out, t, use_ghost)
                if launch is None:
                    continue
                jobs.append(([f, g, h, out], out, launch, lim, (n, ghost, use_ghost, shape, seed)))
                found += 1
                if found == 3:
                    break
            assert found >= 1, (n, ghost, use_ghost)
    finally:
        nh.fused_jit_column()
    with tempfile.TemporaryDirectory() as tmp:
        libs = build([j[2][0] for j in jobs], tmp)
        for so, (fields, out, launch, lim, case) in zip(libs, jobs):
            try:
                audit(so, fields, out, launch, lim)
            except AssertionError as e:
                raise AssertionError(f"{case}: {e}") from None


def test_random_opcode_trees_stay_in_bounds():
    """The random trees of the GPU test that compares compiled kernels with the
    interpreter (every opcode, masks, mask shifts, edgeless stencils, fields with
    two ghost layers): those that run on the column schedule, audited.  One of
    them, an interpolant that reads x and x + 1 only, once loaded the value left
    of a wave's first column for every family, one element before a field's
    allocation in the box's first block."""
    from test_fused_source import random_tree
    n = (13, 9, 11)
    L = nh.Layout.build(n, 2)
    rng = np.random.default_rng(5)
    f, g = host_field(L, "f", rng), host_field(L, "g", rng)
    out = nh.Field.empty(L, "cpu", name="out")
    db = nh.build_stencils(*n, 1.0, 1.0, 1.0)
    m = nh.Mask(L, rng.integers(-1, 5, (40, 3)))
    stencils = [db[("Interpolant", "SVol", "XSurfY")], db[("GradientY", "SVol", "SVol")],
                nh.Stencil([(1, 0, 0)], [2.0], kind=_abi.STENCIL_EDGELESS)]
    r = random.Random(3)
    trees = [random_tree(r, [f, g], stencils, [m, nh.mask_shift(m, (0, 1, 0))], 4) for _ in range(24)]
    jobs = []
    for t in trees:
        try:
            nh.lower_stats(t, "host")
            lim = nh.assign_limits(out, t, True, where="host")
        except nh.NeboError:
            continue
        launch = None if lim is None else column_launch(out, t, True)
        if launch is not None:
            jobs.append((launch, lim))
    assert len(jobs) >= 3
    with tempfile.TemporaryDirectory() as tmp:
        for so, (launch, lim) in zip(build([j[0][0] for j in jobs], tmp), jobs):
            audit(so, [f, g, out], out, launch, lim)


def test_column_launch_refuses_other_trees():
    n = (6, 5, 4)
    L = nh.Layout.build(n, 1)
    rng = np.random.default_rng(0)
    f, g = host_field(L, "f", rng), host_field(L, "g", rng)
    out = nh.Field.empty(L, "cpu")
    assert column_launch(out, f * g + 1.0, True) is None
    m = nh.Mask(L, [(0, 0, 0)])
    assert column_launch(out, nh.cond(m, f)(g), True) is None


def test_dropped_planes_stay_in_bounds():
    """Kernels whose ring drops the plane below (the chunk's first plane formed
    by a prologue from one load of that plane per family): three Laplacians on
    a box 130 wide with two ghost layers (two columns per lane, ghost-inclusive
    and interior regions) and a transport stage at 66^3 (six fields, face
    velocities on their own layouts), audited like the others."""
    from nebo_hip import workloads as W
    jobs = []
    n = (130, 6, 9)
    L = nh.Layout.build(n, 2)
    rng = np.random.default_rng(31)
    f, g, h = host_field(L, "f", rng), host_field(L, "g", rng), host_field(L, "h", rng)
    out = nh.Field.empty(L, "cpu", name="out")
    db = nh.build_stencils(*n, 1.0, 1.0, 1.0)
    lap = lambda q: sum((db[("Divergence", f"SSurf{a}", "SVol")](db[("Gradient", "SVol", f"SSurf{a}")](q))
                         for a in "YZ"), db[("Divergence", "SSurfX", "SVol")](db[("Gradient", "SVol", "SSurfX")](q)))
    t = lap(f) + lap(g) + lap(h)
    for use_ghost in (True, False):
        lim = nh.assign_limits(out, t, use_ghost, where="host")
        launch = column_launch(out, t, use_ghost)
        assert launch is not None and "double P" in launch[0]
        jobs.append(([f, g, h, out], out, launch, lim))
    m = 66
    Lt, Ls, ops, k, phi, u, dt = W.transport_problem(m)
    K = nh.Field.from_numpy(Lt, k, where="cpu", name="k")
    U = [nh.Field.from_numpy(l, x.reshape(-1), where="cpu", name=f"u{d}") for d, (l, x) in enumerate(zip(Ls, u))]
    PN = nh.Field.from_numpy(Lt, phi, where="cpu", name="phin")
    S = nh.Field.from_numpy(Lt, phi.copy(), where="cpu", name="s")
    lo_, hi_ = W.upwind_ops()
    s1, s2, s3 = W.rk3_stage_exprs(ops, PN, K, U, dt, lo_, hi_)
    st = s2(S)
    out_t = nh.Field.empty(Lt, "cpu", name="out")
    lim = nh.assign_limits(out_t, st, True, where="host")
    launch = column_launch(out_t, st, True)
    assert launch is not None and "double P" in launch[0] and "ld2_(rs_" in launch[0]
    jobs.append(([K, PN, S, out_t] + U, out_t, launch, lim))
    with tempfile.TemporaryDirectory() as tmp:
        libs = build([j[2][0] for j in jobs], tmp)
        for so, (fields, lhs, launch, lim) in zip(libs, jobs):
            audit(so, fields, lhs, launch, lim)
