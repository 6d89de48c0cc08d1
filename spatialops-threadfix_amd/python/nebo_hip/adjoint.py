"""Adjoint of the periodic var-coefficient diffusion step (nebo_diffusion_* of
include/nebo_hip.h) and a torch.autograd.Function over a checkpointed trajectory."""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from .api import _stream
from .fields import Layout


def _desc(layout: Layout, t: torch.Tensor, device=0):
    return layout.desc(t.data_ptr(), device)


def diffusion_step(layout, out, phi, k, coefs, dt, stream=None):
    do, dp, dk = _desc(layout, out), _desc(layout, phi), _desc(layout, k)
    _abi.call("nebo_diffusion_step", C.byref(do), C.byref(dp), C.byref(dk), C.byref(coefs), float(dt),
              _stream(stream))


def diffusion_adjoint_step(layout, lam_out, gk, lam, phi, k, coefs, dt, stream=None):
    dlo, dl, dp, dk = (_desc(layout, t) for t in (lam_out, lam, phi, k))
    _abi.call("nebo_diffusion_adjoint_step", C.byref(dlo), C.c_void_p(gk.data_ptr()), C.byref(dl),
              C.byref(dp), C.byref(dk), C.byref(coefs), float(dt), _stream(stream))


def forward(layout, phi0, k, coefs, dt, n_steps, every, stream=None):
    """Returns (phiT, checkpoints)."""
    np_ = layout.npts
    ckpt = torch.empty((n_steps // every + 1) * np_, dtype=torch.float64, device=phi0.device)
    phiT = torch.empty_like(phi0)
    d0, dT, dk = _desc(layout, phi0), _desc(layout, phiT), _desc(layout, k)
    _abi.call("nebo_diffusion_forward", C.byref(dT), C.byref(d0), C.byref(dk), C.byref(coefs), float(dt),
              int(n_steps), int(every), C.c_void_p(ckpt.data_ptr()), _stream(stream))
    return phiT, ckpt


def backward(layout, lamT, k, coefs, dt, n_steps, every, ckpt, stream=None):
    """Returns (dL/dphi0, dL/dk) for dL/dphiT = lamT."""
    lam0 = torch.empty_like(lamT)
    gk = torch.empty_like(k)
    work = torch.empty((every + 1) * layout.npts, dtype=torch.float64, device=lamT.device)
    d0, dT, dk = _desc(layout, lam0), _desc(layout, lamT), _desc(layout, k)
    _abi.call("nebo_diffusion_backward", C.byref(d0), C.c_void_p(gk.data_ptr()), C.byref(dT), C.byref(dk),
              C.byref(coefs), float(dt), int(n_steps), int(every), C.c_void_p(ckpt.data_ptr()),
              C.c_void_p(work.data_ptr()), _stream(stream))
    return lam0, gk


class DiffusionTrajectory(torch.autograd.Function):
    """phiT = (n_steps explicit-Euler steps of d(phi)/dt = div(k grad phi))(phi0), periodic.

    Forward and backward run on the hand-written HIP kernels; backward recomputes
    each `every`-step segment from its checkpoint."""

    @staticmethod
    def forward(ctx, phi0, k, layout, coefs, dt, n_steps, every):
        phiT, ckpt = forward(layout, phi0.detach().contiguous(), k.detach().contiguous(), coefs, dt,
                             n_steps, every)
        ctx.save_for_backward(k, ckpt)
        ctx.cfg = (layout, coefs, dt, n_steps, every)
        return phiT

    @staticmethod
    def backward(ctx, grad_out):
        k, ckpt = ctx.saved_tensors
        layout, coefs, dt, n_steps, every = ctx.cfg
        lam0, gk = backward(layout, grad_out.contiguous(), k, coefs, dt, n_steps, every, ckpt)
        return lam0, gk, None, None, None, None, None
