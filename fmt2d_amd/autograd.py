"""torch.autograd binding of the receiver times.

``TravelTimes`` holds a device context (``fmt2d_amd.hip.Context``) with the
model, the OBS and the rays (source index, path and receiver position of
each).  Calling it with the velocity grid (nz x nx on the propagation grid's
nodes, the layout ``Context.model(0)`` returns; its reciprocal with
``param="slowness"``) and the depths of the chosen interfaces returns the
receiver times of the rays: a forward of all OBS and the rays, on the GPU.

Two backwards:

* ``backward="adjoint"`` (default): the adjoint-state gradient of the
  solver's own discrete times (``Context.adjoint_gradient``,
  fmt2d_adjoint_gradient): a reverse sweep over the updates the forward
  kept, stage by stage, onto every grid node's velocity and every interface
  node's depth.  It is the derivative of the times the forward returns
  (grid velocities held fixed when a depth moves), wherever the times are
  differentiable: where two updates tie (coincident grid and interface
  nodes on the sea surface) it follows the one the forward kept.
* ``backward="rays"``: the adjoint of the linearisation the reference
  inverts with, the Frechet rows of ``compute_frechet_for_ray2d``
  (src/inversion2d.f90:4548-4787), ray paths held fixed.  A velocity node's
  row entry is d t / d s for its slowness s = 1/v, so the gradient with
  respect to the velocity is -(G^T g) / v^2; an interface node's entry is
  d t / d z.  ``Context.frechet_apply`` forms G^T g on the device, each
  column's sum in the order of the reference's ``sparse_ATx``.  Nodes that
  are not inversion parameters (the water, interfaces not inverted) get no
  gradient.

Rays that fail have NaN times and pass no gradient.  The backward needs the
fields and rays of its own forward.  When the context has run another
forward since (a second call, a finite-difference probe), the backward
first repeats its forward with the saved inputs; the forward is
deterministic, so the fields and rays are the same.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch


class TravelTimes:
    """Differentiable receiver times of a fixed set of rays."""

    def __init__(self, ctx, obs_x: Sequence[float], src, path, rec_x, rec_z, invert_velocity: bool = True,
                 ifaces: Sequence[int] = (), backward: str = "adjoint", param: str = "velocity"):
        if backward not in ("adjoint", "rays"):
            raise ValueError("backward must be 'adjoint' or 'rays'")
        if param not in ("velocity", "slowness"):
            raise ValueError("param must be 'velocity' or 'slowness'")
        self.ctx = ctx
        self.obs_x = [float(x) for x in obs_x]
        self.src = np.ascontiguousarray(src, dtype=np.int32)
        self.path = np.ascontiguousarray(path, dtype=np.int32)
        self.rec_x = np.ascontiguousarray(rec_x, dtype=np.float64)
        self.rec_z = np.ascontiguousarray(rec_z, dtype=np.float64)
        self.invert_velocity = bool(invert_velocity)
        self.ifaces = [int(k) for k in ifaces]
        for k in self.ifaces:
            if not 1 <= k <= 5:
                raise ValueError("interface ids are 1..5")
        self.backward = backward
        self.param = param
        self.generation = 0
        self.last_stats = None
        if backward == "rays":
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

    def __call__(self, model: torch.Tensor, *depths: torch.Tensor) -> torch.Tensor:
        if len(depths) != len(self.ifaces):
            raise ValueError(f"expected depths of {len(self.ifaces)} interface(s), got {len(depths)}")
        return _TravelTimesFn.apply(model, self, *depths)

    def velocity_of(self, model: np.ndarray) -> np.ndarray:
        return 1.0 / model if self.param == "slowness" else model

    def _forward(self, model: np.ndarray, depths) -> np.ndarray:
        self.ctx.model_put(0, self.velocity_of(model))
        for k, z in zip(self.ifaces, depths):
            self.ctx.model_put(k, z)
        self.ctx.forward(self.obs_x)
        t, valid = self.ctx.trace(self.src, self.path, self.rec_x, self.rec_z)
        self.generation += 1
        return np.where(valid, t, np.nan), np.asarray(valid, dtype=bool)

    def _gradients(self, g: np.ndarray, v: np.ndarray):
        """(d J / d velocity (nz, nx), [d J / d depth of each interface])."""
        if self.backward == "adjoint":
            gv, gz, st = self.ctx.adjoint_gradient(g)
            self.last_stats = st
            if not self.invert_velocity:
                gv = np.zeros_like(gv)
            return gv, [gz[k - 1] for k in self.ifaces]
        y = np.asarray(self.ctx.frechet_apply(g, transpose=True))
        idx = self.node_param
        sel = (idx > 0) & (idx <= self.n_vel)
        gv = np.zeros(v.shape)
        gv[sel] = -y[idx[sel] - 1] / (v[sel] * v[sel])
        return gv, [y[sl] for sl in self.iface_slices]


class _TravelTimesFn(torch.autograd.Function):

    @staticmethod
    def forward(fctx, model_t, tt: TravelTimes, *depths):
        m = model_t.detach().to("cpu", torch.float64).numpy()
        zs = [z.detach().to("cpu", torch.float64).numpy() for z in depths]
        t, valid = tt._forward(m, zs)
        fctx.tt = tt
        fctx.generation = tt.generation
        fctx.valid = valid
        fctx.save_for_backward(model_t, *depths)
        return torch.as_tensor(t, dtype=model_t.dtype, device=model_t.device)

    @staticmethod
    def backward(fctx, grad_t):
        tt: TravelTimes = fctx.tt
        model_t, *depths = fctx.saved_tensors
        m = model_t.detach().to("cpu", torch.float64).numpy()
        if tt.generation != fctx.generation:
            tt._forward(m, [z.detach().to("cpu", torch.float64).numpy() for z in depths])
        g = grad_t.detach().to("cpu", torch.float64).numpy()
        g = np.where(fctx.valid, g, 0.0)
        v = tt.velocity_of(m)
        gv, gzs = tt._gradients(g, v)
        grad_m = None
        if fctx.needs_input_grad[0]:
            gm = -gv * v * v if tt.param == "slowness" else gv
            grad_m = torch.as_tensor(gm, dtype=model_t.dtype, device=model_t.device)
        grads_z = []
        for k, (z, gz) in enumerate(zip(depths, gzs)):
            if fctx.needs_input_grad[2 + k]:
                grads_z.append(torch.as_tensor(gz, dtype=z.dtype, device=z.device))
            else:
                grads_z.append(None)
        return (grad_m, None, *grads_z)
