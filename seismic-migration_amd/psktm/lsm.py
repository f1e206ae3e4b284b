"""Least-squares PSTM: autograd wrappers and conjugate gradients.

The forward operator of least-squares migration is Kirchhoff demigration
A: image -> traces (the adjoint of the migration kernel); its adjoint A' is
migration. Both are the HIP kernels behind ``psktm_demigrate`` and
``psktm_migrate`` on one plan, so they share the discrete traveltime,
interpolation weights, aperture and amplitude weight.

* :class:`Migrate` / :class:`Demigrate` - ``torch.autograd.Function`` pairs
  whose backward is the other kernel (a test wrapper: gradients of any loss
  through them are exact adjoints up to FP32 rounding).
* ``gained=True`` selects the gained operator for both: migration of
  gained traces (:func:`trace_gain`, then :func:`migrate` with
  ``gained=True``) and its adjoint (:func:`demigrate` with ``gained=True``):
  the weight's sqrt(t) acts on the trace samples, one ``v_sqrt_f32`` fewer
  per update in each kernel.
* :func:`cgls` - conjugate gradients on the normal equations,
  min ||A m - d||^2, with every inner product accumulated in FP64 on the GPU
  and the step lengths alpha = gamma/delta, beta = gamma'/gamma formed on the
  device from those FP64 scalars (``psktm_dot_f64_ws``, in a fixed order, so
  equal vectors give equal scalars; ``psktm_axpy_ratio``,
  ``psktm_xpby_ratio``); an iteration never waits for the host and allocates
  nothing, so it can be captured in a HIP graph.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .api import (Plan, _need, _ptr, _stream, demigrate, dot_f64, dot_workspace, migrate,
                  trace_gain)

_p = ctypes.c_void_p


import torch


def _migrate_op(plan, traces, vel, image=None, gained=False, work=None):
    """L(traces): migration, or migration of the gained traces."""
    if not gained:
        return migrate(plan, traces, vel, image)
    g = trace_gain(traces, plan.survey.dt, out=work)
    return migrate(plan, g, vel, image, gained=True)


class Migrate(torch.autograd.Function):
    """image = L(traces) (migration); backward: grad_traces = L'(grad_image)."""

    @staticmethod
    def forward(ctx, traces, plan, vel, gained=False):
        ctx.plan, ctx.vel, ctx.gained = plan, vel, gained
        return _migrate_op(plan, traces.contiguous(), vel, gained=gained)

    @staticmethod
    def backward(ctx, grad_image):
        return (demigrate(ctx.plan, grad_image.contiguous(), ctx.vel, gained=ctx.gained),
                None, None, None)


class Demigrate(torch.autograd.Function):
    """traces = L'(image) (demigration); backward: grad_image = L(grad_traces)."""

    @staticmethod
    def forward(ctx, image, plan, vel, gained=False):
        ctx.plan, ctx.vel, ctx.gained = plan, vel, gained
        return demigrate(plan, image.contiguous(), vel, gained=gained)

    @staticmethod
    def backward(ctx, grad_traces):
        return (_migrate_op(ctx.plan, grad_traces.contiguous(), ctx.vel, gained=ctx.gained),
                None, None, None)


def _axpy_ratio(y, x, num, den, sign, stream=None):
    L = _lib.load()
    _lib.check(L.psktm_axpy_ratio(_p(_ptr(y)), _p(_ptr(x)), ctypes.c_size_t(y.numel()),
                                  _p(_ptr(num)), _p(_ptr(den)), float(sign), _stream(stream)))


def _xpby_ratio(p, r, num, den, stream=None):
    L = _lib.load()
    _lib.check(L.psktm_xpby_ratio(_p(_ptr(p)), _p(_ptr(r)), ctypes.c_size_t(p.numel()),
                                  _p(_ptr(num)), _p(_ptr(den)), _stream(stream)))


class CGLS:
    """Conjugate gradients for min ||A m - d||^2 with A = demigration.

    State lives on the device: m (image), r (data residual), s = A' r,
    p (search direction), q = A p, and FP64 scalars gamma = <s,s>,
    delta = <q,q>, rr = <r,r>. :meth:`step` enqueues one iteration on the
    current stream without synchronising.
    """

    def __init__(self, plan: Plan, vel, data, m0=None, gained=False):
        _need(data, torch.float32, "data", plan.survey.n_traces * plan.survey.n_samples)
        self.plan, self.vel, self.gained = plan, vel, gained
        dev = data.device
        g = plan.grid
        self.m = torch.zeros((g.nx, g.ny, g.nt), dtype=torch.float32, device=dev) \
            if m0 is None else m0.clone()
        self.r = data.clone()
        self.work = torch.empty_like(data) if gained else None     # gained residual
        if m0 is not None:
            self.r -= demigrate(plan, self.m, vel, gained=gained)
        self.s = _migrate_op(plan, self.r, vel, gained=gained, work=self.work)
        self.p = self.s.clone()
        self.q = torch.empty_like(data)
        self.ws = dot_workspace(dev)
        self.gamma = self._dot(self.s, self.s)
        self.gamma_new = torch.empty_like(self.gamma)
        self.delta = torch.empty_like(self.gamma)
        self.rr = self._dot(self.r, self.r)

    def _dot(self, a, b, out=None):
        return dot_f64(a.reshape(-1), b.reshape(-1), out, workspace=self.ws)

    def step(self):
        plan, vel = self.plan, self.vel
        demigrate(plan, self.p, vel, self.q, gained=self.gained)           # q = A p
        self._dot(self.q, self.q, self.delta)                              # delta = <q,q>
        _axpy_ratio(self.m, self.p, self.gamma, self.delta, +1.0)          # m += alpha p
        _axpy_ratio(self.r, self.q, self.gamma, self.delta, -1.0)          # r -= alpha q
        _migrate_op(plan, self.r, vel, self.s, self.gained, self.work)     # s = A' r
        self._dot(self.s, self.s, self.gamma_new)                          # gamma' = <s,s>
        _xpby_ratio(self.p, self.s, self.gamma_new, self.gamma)            # p = s + beta p
        self.gamma, self.gamma_new = self.gamma_new, self.gamma
        self._dot(self.r, self.r, self.rr)                                 # ||r||^2


def cgls(plan: Plan, vel, data, iters: int = 20, m0=None, history: bool = True,
         gained: bool = False):
    """Run ``iters`` CGLS iterations; returns (image, [||r_k||^2 ...])."""
    cg = CGLS(plan, vel, data, m0, gained)
    hist = [cg.rr.clone()] if history else []
    for _ in range(iters):
        cg.step()
        if history:
            hist.append(cg.rr.clone())
    torch.cuda.synchronize()
    return cg.m, [float(h.item()) for h in hist]
