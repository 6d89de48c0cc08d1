"""The benchmark workloads written as Nebo expressions (shared by tests and bench).

heat_*      : explicit Euler heat equation with zero-Dirichlet ghosts set by a mask
              and max/sum reductions each step (examples/heat_equation.cpp pattern).
transport_* : scalar transport, upwinded convective flux selected with cond on the
              sign of a face velocity plus the diffusive flux div(k grad phi),
              SSP-RK3 in time; every RK stage is one fused assignment.
"""
from __future__ import annotations

import numpy as np

import ctypes as C

from . import _abi, inputs
from .api import _stream
from .expr import E, Stencil, cond, serialize
from .fields import Field, Layout, Mask
from .ops import basic_op_types, build_stencils


# --------------------------------------------------------------------------
# heat equation
# --------------------------------------------------------------------------

def heat_initial(n, seed=1, sigma=0.1, noise=1e-3):
    """Gaussian (sigma 0.1, centred) plus seeded uniform noise; zero ghosts."""
    x = inputs.cell_centres(n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    g = np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2) / (2 * sigma * sigma))
    g = g + noise * inputs.uniform(seed, g.size, -1.0, 1.0).reshape(g.shape)
    g[0, :, :] = g[-1, :, :] = 0.0
    g[:, 0, :] = g[:, -1, :] = 0.0
    g[:, :, 0] = g[:, :, -1] = 0.0
    return g.reshape(-1)


def ghost_points(n):
    """Every ghost cell of an n^3 interior with one ghost layer (interior-relative)."""
    r = np.arange(-1, n + 1)
    Z, Y, X = np.meshgrid(r, r, r, indexing="ij")
    on = (X == -1) | (X == n) | (Y == -1) | (Y == n) | (Z == -1) | (Z == n)
    return np.stack([X[on], Y[on], Z[on]], 1)


def heat_problem(n=64, alpha=1.0):
    L = Layout.build((n, n, n), 1)
    db = build_stencils(n, n, n, 1.0, 1.0, 1.0)
    ops = basic_op_types(db)
    dx = 1.0 / n
    dt = 0.1 * dx * dx / alpha
    mask = Mask(L, ghost_points(n))
    return L, ops, dt * alpha, mask


def heat_step_exprs(ops, phi, s, mask):
    """(phi_new <<= phi + s*lap(phi), phi_new <<= cond(ghosts, 0)(phi_new))."""
    lap = (ops["DivX"](ops["GradX"](phi)) + ops["DivY"](ops["GradY"](phi))
           + ops["DivZ"](ops["GradZ"](phi)))
    return phi + s * lap, (lambda out: cond(mask, 0.0)(out))


class HeatStepper:
    """The heat-equation time step of examples/heat_equation.cpp:216-242 through
    the C-ABI, with the expression descriptors serialised once (as a time loop
    keeps them): phi' <<= phi + s*lap(phi) (the Laplacian + Euler kernel),
    phi' <<= cond(ghost mask, 0)(phi'), then field_max(phi') and field_sum(phi')
    written to device memory (red[2i], red[2i + 1]) without a host
    synchronisation.  Step i reads block i % 2 and writes the other, so an even
    number of steps ends in block 0.  Every call allocates nothing, so a run of
    steps on one stream can be captured in a hipGraph once a step has run on that
    stream (the first reduction on a stream creates its working buffer)."""

    def __init__(self, n=64, slots=100):
        import torch
        L, ops, s, mask = heat_problem(n)
        mask.add_consumer(0)
        self.n, self.layout = n, L
        self.phi0 = heat_initial(n)
        self.fields = [Field.from_numpy(L, self.phi0, where="gpu"), Field.from_numpy(L, self.phi0, where="gpu")]
        self.red = torch.zeros(2 * slots, dtype=torch.float64, device="cuda")
        self.slots = slots
        self._plans = []
        for src in (0, 1):
            e1, cond_of = heat_step_exprs(ops, self.fields[src], s, mask)
            dst = self.fields[1 - src]
            d, s1, s2, sf = dst.desc("gpu"), serialize(e1, "gpu"), serialize(cond_of(dst), "gpu"), serialize(E(dst), "gpu")
            # ctypes arguments made once: a time loop only passes them again
            self._plans.append((C.byref(d), s1.ref, s2.ref, sf.ref, (d, s1, s2, sf)))
        self._red0 = self.red.data_ptr()
        lib = _abi.lib()
        self._assign, self._reduce = lib.nebo_assign, lib.nebo_reduce_async
        self._keep = (ops, mask)

    def reset(self):
        import torch
        self.fields[0].gpu.copy_(torch.from_numpy(self.phi0))

    def step(self, i, stream=None):
        sp = _stream(stream)
        d, e1, e2, ef, _ = self._plans[i & 1]
        slot = self._red0 + 16 * (i % self.slots)   # red[2 * slot], then red[2 * slot + 1]
        rc = (self._assign(d, e1, 1, sp) or self._assign(d, e2, 1, sp)
              or self._reduce(ef, _abi.RED_MAX, 1, slot, sp) or self._reduce(ef, _abi.RED_SUM, 1, slot + 8, sp))
        if rc:
            _abi.check(rc)


# --------------------------------------------------------------------------
# scalar transport
# --------------------------------------------------------------------------

def transport_velocity(n, seed=7, modes=4):
    """A divergence-free face velocity u = curl(A) on the staggered grid, scaled so
    max |u| = 1.  A is a sum of seeded Fourier modes (edge-centred), so the discrete
    divergence of the face velocity is zero up to rounding."""
    h = 1.0 / n
    idx = np.arange(-1, n + 1, dtype=np.float64)
    rng_k = (inputs.splitmix64(seed, 3 * modes) % np.uint64(4)).astype(np.int64).reshape(modes, 3) + 1
    amp = inputs.normal(seed + 1, 3 * modes).reshape(modes, 3)
    ph = inputs.uniform(seed + 2, 3 * modes, 0.0, 2 * np.pi).reshape(modes, 3)

    def A(comp, sx, sy, sz):
        # component comp of the potential at window points shifted by (sx, sy, sz) cells
        z = (idx + sz)[:, None, None] * h
        y = (idx + sy)[None, :, None] * h
        x = (idx + sx)[None, None, :] * h
        out = np.zeros((n + 2,) * 3)
        for m in range(modes):
            kx, ky, kz = rng_k[m]
            out += amp[m, comp] * np.sin(2 * np.pi * (kx * x + ky * y + kz * z) + ph[m, comp])
        return out

    # Cell (i,j,k) has centre ((i+.5)h, (j+.5)h, (k+.5)h); its low x face is at x = i h.
    # Edge potentials: Az at (i h, j h, (k+.5)h), Ay at (i h, (j+.5)h, k h),
    # Ax at ((i+.5)h, j h, k h).  Face velocities are the discrete curl, so the
    # discrete divergence of (ux, uy, uz) cancels term by term.
    Az0, Az_y1, Az_x1 = A(2, 0.0, 0.0, 0.5), A(2, 0.0, 1.0, 0.5), A(2, 1.0, 0.0, 0.5)
    Ay0, Ay_z1, Ay_x1 = A(1, 0.0, 0.5, 0.0), A(1, 0.0, 0.5, 1.0), A(1, 1.0, 0.5, 0.0)
    Ax0, Ax_z1, Ax_y1 = A(0, 0.5, 0.0, 0.0), A(0, 0.5, 0.0, 1.0), A(0, 0.5, 1.0, 0.0)
    ux = ((Az_y1 - Az0) - (Ay_z1 - Ay0)) / h
    uy = ((Ax_z1 - Ax0) - (Az_x1 - Az0)) / h
    uz = ((Ay_x1 - Ay0) - (Ax_y1 - Ax0)) / h
    umax = max(np.abs(ux).max(), np.abs(uy).max(), np.abs(uz).max())
    return [inputs.periodic_fill(u / umax, 1).reshape(-1) for u in (ux, uy, uz)]


def upwind_ops():
    """One-point shift stencils SVol -> SSurf{X,Y,Z}: the value of the cell on the
    low side of a face and of the cell on its high side."""
    lo = [Stencil([(-1, 0, 0)], [1.0], src="SVol", dst="SSurfX"),
          Stencil([(0, -1, 0)], [1.0], src="SVol", dst="SSurfY"),
          Stencil([(0, 0, -1)], [1.0], src="SVol", dst="SSurfZ")]
    hi = [Stencil([(0, 0, 0)], [1.0], src="SVol", dst=d) for d in ("SSurfX", "SSurfY", "SSurfZ")]
    return lo, hi


def transport_rhs(ops, phi, k, u, lo, hi):
    """-(sum_d div_d(u_d * upwind_d(phi))) + sum_d div_d(interp_d(k) * grad_d(phi))."""
    conv = None
    diff = None
    for d, ax in enumerate("XYZ"):
        up = cond(E(u[d]) > 0.0, lo[d](phi))(hi[d](phi))
        c = ops[f"Div{ax}"](up * u[d])
        f = ops[f"Div{ax}"](ops[f"InterpC2F{ax}"](k) * ops[f"Grad{ax}"](phi))
        conv = c if conv is None else conv + c
        diff = f if diff is None else diff + f
    return diff - conv


def rk3_stage_exprs(ops, phin, k, u, dt, lo, hi):
    """The three SSP-RK3 stage right-hand sides given the stage fields:
    s1 <<= phin + dt*R(phin);  s2 <<= 0.75*phin + 0.25*(s1 + dt*R(s1));
    phin' <<= (1/3)*phin + (2/3)*(s2 + dt*R(s2))."""
    def stage1():
        return phin + dt * transport_rhs(ops, phin, k, u, lo, hi)

    def stage2(s1):
        return 0.75 * E(phin) + 0.25 * (s1 + dt * transport_rhs(ops, s1, k, u, lo, hi))

    def stage3(s2):
        return (1.0 / 3.0) * E(phin) + (2.0 / 3.0) * (s2 + dt * transport_rhs(ops, s2, k, u, lo, hi))

    return stage1, stage2, stage3


def transport_problem(n, seed_k=2, seed_phi=3, seed_u=7):
    L = Layout.build((n, n, n), 1)
    Lx, Ly, Lz = L.like("SSurfX"), L.like("SSurfY"), L.like("SSurfZ")
    db = build_stencils(n, n, n, 1.0, 1.0, 1.0)
    ops = basic_op_types(db)
    k = inputs.coefficient_field(n, seed_k, 0.5, 2.0).reshape(-1) * 1e-3
    phi = inputs.fourier_field(n, seed_phi).reshape(-1)
    u = transport_velocity(n, seed_u)
    h = 1.0 / n
    dt = 0.2 * min(h, h * h / (6 * 2e-3))
    return L, (Lx, Ly, Lz), ops, k, phi, u, dt


__all__ = ["heat_initial", "ghost_points", "heat_problem", "heat_step_exprs", "HeatStepper", "transport_velocity",
           "upwind_ops", "transport_rhs", "rk3_stage_exprs", "transport_problem", "Field"]


def transport_problem_gpu(n, device=0, seed_k=2, seed_phi=3, seed_u=7, modes_phi=8, modes_u=4):
    """transport_problem's fields generated on the GPU with torch (same formulas and
    seeds; the device sin differs from numpy's in the last bits, so use this only
    where no CPU result is compared, e.g. the 512^3 benchmark leg)."""
    import torch
    dev = torch.device(f"cuda:{device}")
    L = Layout.build((n, n, n), 1)
    Ls = (L.like("SSurfX"), L.like("SSurfY"), L.like("SSurfZ"))
    db = build_stencils(n, n, n, 1.0, 1.0, 1.0)
    ops = basic_op_types(db)
    h = 1.0 / n
    idx = torch.arange(-1, n + 1, dtype=torch.float64, device=dev)

    def periodic(a):
        a[0], a[-1] = a[-2].clone(), a[1].clone()
        a[:, 0], a[:, -1] = a[:, -2].clone(), a[:, 1].clone()
        a[:, :, 0], a[:, :, -1] = a[:, :, -2].clone(), a[:, :, 1].clone()
        return a

    # k ~ U[0.5, 2) * 1e-3 on the interior (seeded on the host, n^3 draws), periodic ghosts
    k = torch.zeros((n + 2,) * 3, dtype=torch.float64, device=dev)
    k[1:-1, 1:-1, 1:-1] = torch.from_numpy(inputs.uniform(seed_k, n ** 3, 0.5, 2.0).reshape(n, n, n)).to(dev)
    k = periodic(k * 1e-3)
    # phi: 8 Fourier modes at cell centres
    kk = (inputs.splitmix64(seed_phi, 3 * modes_phi) % np.uint64(8)).astype(np.int64).reshape(modes_phi, 3) + 1
    amp = inputs.normal(seed_phi + 1, modes_phi)
    th = inputs.uniform(seed_phi + 2, modes_phi, 0.0, 2 * np.pi)
    xc = (idx + 0.5) * h
    phi = torch.zeros((n + 2,) * 3, dtype=torch.float64, device=dev)
    for m in range(modes_phi):
        phi += float(amp[m]) * torch.sin(2 * np.pi * (int(kk[m, 2]) * xc[:, None, None] + int(kk[m, 1]) * xc[None, :, None]
                                                     + int(kk[m, 0]) * xc[None, None, :]) + float(th[m]))
    # velocity: the discrete curl of edge-centred potentials (as transport_velocity)
    rk = (inputs.splitmix64(seed_u, 3 * modes_u) % np.uint64(4)).astype(np.int64).reshape(modes_u, 3) + 1
    ua = inputs.normal(seed_u + 1, 3 * modes_u).reshape(modes_u, 3)
    up = inputs.uniform(seed_u + 2, 3 * modes_u, 0.0, 2 * np.pi).reshape(modes_u, 3)

    def A(comp, sx, sy, sz):
        z = (idx + sz)[:, None, None] * h
        y = (idx + sy)[None, :, None] * h
        x = (idx + sx)[None, None, :] * h
        out = torch.zeros((n + 2,) * 3, dtype=torch.float64, device=dev)
        for m in range(modes_u):
            out += float(ua[m, comp]) * torch.sin(2 * np.pi * (int(rk[m, 0]) * x + int(rk[m, 1]) * y
                                                              + int(rk[m, 2]) * z) + float(up[m, comp]))
        return out

    ux = ((A(2, 0, 1, .5) - A(2, 0, 0, .5)) - (A(1, 0, .5, 1) - A(1, 0, .5, 0))) / h
    uy = ((A(0, .5, 0, 1) - A(0, .5, 0, 0)) - (A(2, 1, 0, .5) - A(2, 0, 0, .5))) / h
    uz = ((A(1, 1, .5, 0) - A(1, 0, .5, 0)) - (A(0, .5, 1, 0) - A(0, .5, 0, 0))) / h
    umax = max(float(ux.abs().max()), float(uy.abs().max()), float(uz.abs().max()))
    u = [periodic(v / umax).reshape(-1) for v in (ux, uy, uz)]
    dt = 0.2 * min(h, h * h / (6 * 2e-3))
    return L, Ls, ops, k.reshape(-1), periodic(phi).reshape(-1), u, dt
