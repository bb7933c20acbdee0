"""torch.autograd binding of the receiver times.

``TravelTimes`` holds a device context (``fmt2d_amd.hip.Context``) with the
model, the OBS and the rays (source index, path and receiver position of
each).  Calling it with the velocity grid (nz x nx on the propagation grid's
nodes, the layout ``Context.model(0)`` returns) and the depths of the
inverted interfaces returns the receiver times of the rays: a forward of all
OBS and the rays, on the GPU.

The backward is the adjoint of the linearisation the reference inverts
with: the Frechet rows of ``compute_frechet_for_ray2d``
(src/inversion2d.f90:4548-4787), ray paths held fixed.  A velocity node's
row entry is d t / d s for its slowness s = 1/v, so the gradient with
respect to the velocity is -(G^T g) / v^2; an interface node's entry is
d t / d z.  ``Context.frechet_apply`` forms G^T g on the device, each
column's sum in the order of the reference's ``sparse_ATx``.  Nodes that
are not inversion parameters (the water, interfaces not inverted) get no
gradient.  Rays that fail have NaN times and pass no gradient.

The backward needs the rays of its own forward.  When the context has run
another forward since (a second call, a finite-difference probe), the
backward first repeats its forward with the saved inputs; the forward is
deterministic, so the rays are the same.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch


class TravelTimes:
    """Differentiable receiver times of a fixed set of rays."""

    def __init__(self, ctx, obs_x: Sequence[float], src, path, rec_x, rec_z, invert_velocity: bool = True,
                 ifaces: Sequence[int] = ()):
        self.ctx = ctx
        self.obs_x = [float(x) for x in obs_x]
        self.src = np.ascontiguousarray(src, dtype=np.int32)
        self.path = np.ascontiguousarray(path, dtype=np.int32)
        self.rec_x = np.ascontiguousarray(rec_x, dtype=np.float64)
        self.rec_z = np.ascontiguousarray(rec_z, dtype=np.float64)
        self.invert_velocity = bool(invert_velocity)
        self.ifaces = [int(k) for k in ifaces]
        ctx.set_inversion(self.invert_velocity, self.ifaces)
        self.n_vel, self.n_par = ctx.inversion_info()
        self.node_param = np.asarray(ctx.param_index())
        sizes = [len(ctx.model(k)) for k in self.ifaces]
        self.iface_slices = []
        start = self.n_vel
        for n in sizes:
            self.iface_slices.append(slice(start, start + n))
            start += n
        if start != self.n_par:
            raise ValueError("interface parameters do not add up to the parameter count")
        self.generation = 0

    def __call__(self, velocity: torch.Tensor, *depths: torch.Tensor) -> torch.Tensor:
        if len(depths) != len(self.ifaces):
            raise ValueError(f"expected depths of {len(self.ifaces)} interface(s), got {len(depths)}")
        return _TravelTimesFn.apply(velocity, self, *depths)

    def _forward(self, velocity: np.ndarray, depths) -> np.ndarray:
        self.ctx.model_put(0, velocity)
        for k, z in zip(self.ifaces, depths):
            self.ctx.model_put(k, z)
        self.ctx.forward(self.obs_x)
        t, valid = self.ctx.trace(self.src, self.path, self.rec_x, self.rec_z)
        self.generation += 1
        return np.where(valid, t, np.nan), np.asarray(valid, dtype=bool)

    def _adjoint(self, g: np.ndarray) -> np.ndarray:
        return np.asarray(self.ctx.frechet_apply(g, transpose=True))


class _TravelTimesFn(torch.autograd.Function):

    @staticmethod
    def forward(fctx, velocity, model: TravelTimes, *depths):
        v = velocity.detach().to("cpu", torch.float64).numpy()
        zs = [z.detach().to("cpu", torch.float64).numpy() for z in depths]
        t, valid = model._forward(v, zs)
        fctx.model = model
        fctx.generation = model.generation
        fctx.valid = valid
        fctx.save_for_backward(velocity, *depths)
        return torch.as_tensor(t, dtype=velocity.dtype, device=velocity.device)

    @staticmethod
    def backward(fctx, grad_t):
        model: TravelTimes = fctx.model
        velocity, *depths = fctx.saved_tensors
        if model.generation != fctx.generation:
            model._forward(velocity.detach().to("cpu", torch.float64).numpy(),
                           [z.detach().to("cpu", torch.float64).numpy() for z in depths])
        g = grad_t.detach().to("cpu", torch.float64).numpy()
        g = np.where(fctx.valid, g, 0.0)
        y = model._adjoint(g)
        grad_v = None
        if fctx.needs_input_grad[0]:
            v = velocity.detach().to("cpu", torch.float64).numpy()
            idx = model.node_param
            sel = (idx > 0) & (idx <= model.n_vel)
            gv = np.zeros(v.shape)
            gv[sel] = -y[idx[sel] - 1] / (v[sel] * v[sel])
            grad_v = torch.as_tensor(gv, dtype=velocity.dtype, device=velocity.device)
        grads_z = []
        for k, (z, sl) in enumerate(zip(depths, model.iface_slices)):
            if fctx.needs_input_grad[2 + k]:
                grads_z.append(torch.as_tensor(y[sl], dtype=z.dtype, device=z.device))
            else:
                grads_z.append(None)
        return (grad_v, None, *grads_z)
