"""Scalar transport chain (nebo_transport_* of include/nebo_hip.h): one SSP-RK3
stage of R(s) = div(interp(k)*grad(s)) - div(u*upwind(s)) on the hand-written
kernel, its transpose, the checkpointed forward / reverse sweeps, and a
torch.autograd.Function over a whole trajectory (gradients in phi0 and k)."""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from .api import _stream
from .fields import Field, Layout


def _faces(face_layouts, us, device=0):
    """(FieldT * 3) of the three face-velocity tensors (or Fields) on the GPU."""
    arr = (_abi.FieldT * 3)()
    for d in range(3):
        u = us[d]
        arr[d] = u.desc("gpu") if isinstance(u, Field) else face_layouts[d].desc(u.data_ptr(), device)
    return arr


def _desc(layout: Layout, t, device=0):
    return t.desc("gpu") if isinstance(t, Field) else layout.desc(t.data_ptr(), device)


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


SCHED_AUTO, SCHED_DIRECT, SCHED_COLUMN, SCHED_COLUMN_PF1 = 0, 1, 2, 3


def tune(schedule=SCHED_AUTO, zchunk=0):
    """Schedule of the stage kernel (chosen by box size, direct, column tiles
    with loads 2 or 1 planes ahead) and planes per workgroup (0: its default)."""
    _abi.call("nebo_transport_tune", int(schedule), int(zchunk))


def schedule(layout, device=0):
    """(schedule, z chunk) a stage over a field of this layout on device would
    use (the query reads only the descriptor's sizes)."""
    d = layout.desc(None, device)
    sc, zc = C.c_int32(), C.c_int32()
    _abi.call("nebo_transport_schedule", C.byref(d), C.byref(sc), C.byref(zc))
    return sc.value, zc.value


def stage(layout, out, phin, s, k, us, coefs, dt, stage, face_layouts=None, stream=None):
    """out <<= stage `stage` (1, 2, 3) of SSP-RK3, then out's periodic ghosts.
    Operands are Fields or flat tensors of `layout` (faces: of `face_layouts`)."""
    do, dn, dk = _desc(layout, out), _desc(layout, phin), _desc(layout, k)
    ds = C.byref(_desc(layout, s)) if s is not None else None
    _abi.call("nebo_transport_stage", C.byref(do), C.byref(dn), ds, C.byref(dk), _faces(face_layouts, us),
              C.byref(coefs), float(dt), int(stage), _stream(stream))


def adjoint_stage(layout, out, acc_out, gk, a, acc_in, s, k, us, coefs, dt, stage, face_layouts=None,
                  stream=None):
    """Transpose of one stage (see nebo_transport_adjoint_stage); acc_out, acc_in
    and gk are flat tensors of the field's size."""
    do, da, ds, dk = (_desc(layout, t) for t in (out, a, s, k))
    _abi.call("nebo_transport_adjoint_stage", C.byref(do), _ptr(acc_out), _ptr(gk), C.byref(da), _ptr(acc_in),
              C.byref(ds), C.byref(dk), _faces(face_layouts, us), C.byref(coefs), float(dt), int(stage),
              _stream(stream))


def forward(layout, phi0, k, us, coefs, dt, n_steps, every, face_layouts=None, stream=None):
    """Returns (phiT, checkpoints)."""
    np_ = layout.npts
    ckpt = torch.empty((n_steps // every + 1) * np_, dtype=torch.float64, device=phi0.device)
    phiT = torch.empty_like(phi0)
    d0, dT, dk = _desc(layout, phi0), _desc(layout, phiT), _desc(layout, k)
    _abi.call("nebo_transport_forward", C.byref(dT), C.byref(d0), C.byref(dk), _faces(face_layouts, us),
              C.byref(coefs), float(dt), int(n_steps), int(every), _ptr(ckpt), _stream(stream))
    return phiT, ckpt


def backward(layout, lamT, k, us, coefs, dt, n_steps, every, ckpt, face_layouts=None, stream=None):
    """Returns (dL/dphi0, dL/dk) for dL/dphiT = lamT."""
    lam0 = torch.empty_like(lamT)
    gk = torch.empty_like(k)
    work = torch.empty((3 * every + 5) * layout.npts, dtype=torch.float64, device=lamT.device)
    d0, dT, dk = _desc(layout, lam0), _desc(layout, lamT), _desc(layout, k)
    _abi.call("nebo_transport_backward", C.byref(d0), _ptr(gk), C.byref(dT), C.byref(dk),
              _faces(face_layouts, us), C.byref(coefs), float(dt), int(n_steps), int(every), _ptr(ckpt),
              _ptr(work), _stream(stream))
    return lam0, gk


class TransportTrajectory(torch.autograd.Function):
    """phiT = (n_steps SSP-RK3 steps of d(phi)/dt = R(phi))(phi0), periodic; the face
    velocity is fixed.  Backward recomputes each `every`-step segment from its
    checkpoint and sweeps the transposed stages on the hand-written kernels."""

    @staticmethod
    def forward(ctx, phi0, k, layout, face_layouts, us, coefs, dt, n_steps, every):
        phiT, ckpt = forward(layout, phi0.detach().contiguous(), k.detach().contiguous(), us, coefs, dt,
                             n_steps, every, face_layouts)
        ctx.save_for_backward(k, ckpt)
        ctx.cfg = (layout, face_layouts, us, coefs, dt, n_steps, every)
        return phiT

    @staticmethod
    def backward(ctx, grad_out):
        k, ckpt = ctx.saved_tensors
        layout, face_layouts, us, coefs, dt, n_steps, every = ctx.cfg
        lam0, gk = backward(layout, grad_out.contiguous(), k, us, coefs, dt, n_steps, every, ckpt, face_layouts)
        return lam0, gk, None, None, None, None, None, None, None
