"""The autograd plumbing of fmt2d_amd.autograd without a GPU: a stand-in
context whose "receiver times" are exactly linear in the slowness of the
parameter nodes and the interface depths, t = Gv (1/v) + Gz z, with the
same context methods as fmt2d_amd.hip.Context.  torch.autograd.gradcheck
then compares the backward (adjoint applied per parameter, mapped to grid
nodes, chain rule through s = 1/v) with finite differences of that
forward.  A failing ray (NaN time) passes no gradient."""
import numpy as np
import torch

from fmt2d_amd.autograd import TravelTimes

NZ, NX, NRAY, NIF = 4, 5, 7, 5


class LinearCtx:
    def __init__(self, seed=0):
        rng = np.random.default_rng(seed)
        self.idx = np.zeros((NZ, NX), dtype=np.int32)
        k = 0
        for iz in range(1, NZ):          # row 0 is water: no parameters
            for ix in range(NX):
                k += 1
                self.idx[iz, ix] = k
        self.n_vel = k
        self.G = rng.normal(size=(NRAY, self.n_vel + NIF))
        self.v = np.full((NZ, NX), 3.0)
        self.z = np.zeros(NIF)
        self.bad = 3                     # this ray fails

    def set_inversion(self, invert_velocity, ifaces):
        assert invert_velocity and ifaces == [4]

    def inversion_info(self):
        return self.n_vel, self.n_vel + NIF

    def param_index(self):
        return self.idx

    def model(self, what=0):
        return self.v if what == 0 else self.z

    def model_put(self, what, values):
        if what == 0:
            self.v = np.asarray(values, dtype=float).reshape(NZ, NX).copy()
        else:
            self.z = np.asarray(values, dtype=float).copy()

    def forward(self, obs_x):
        pass

    def trace(self, src, path, rx, rz):
        order = np.argsort(self.idx[self.idx > 0])      # parameter order
        s = np.empty(self.n_vel)
        s[order] = 1.0 / self.v[self.idx > 0]
        m = np.concatenate([s, self.z])
        t = self.G @ m
        valid = np.ones(NRAY, dtype=bool)
        valid[self.bad] = False
        return t, valid

    def frechet_apply(self, x, transpose=False):
        assert transpose
        return self.G.T @ np.asarray(x)

    def adjoint_gradient(self, g):
        """The stand-in's exact gradient in fmt2d_adjoint_gradient's layout:
        d J / d velocity of every grid node, d J / d depth of interfaces 1..5."""
        y = self.G.T @ np.asarray(g)
        order = np.argsort(self.idx[self.idx > 0])
        gs = np.empty(self.n_vel)
        gs[:] = y[:self.n_vel][order]
        gv = np.zeros((NZ, NX))
        gv[self.idx > 0] = -gs / self.v[self.idx > 0] ** 2
        gz = np.zeros((5, NX))
        gz[3] = y[self.n_vel:]
        return gv, gz, None


def make(backward="rays", param="velocity"):
    ctx = LinearCtx()
    tt = TravelTimes(ctx, [10.0], np.zeros(NRAY), np.ones(NRAY), np.arange(NRAY, dtype=float), np.zeros(NRAY),
                     invert_velocity=True, ifaces=[4], backward=backward, param=param)
    return ctx, tt


def test_gradcheck_velocity_and_depths():
    ctx, tt = make()
    rng = np.random.default_rng(1)
    v = torch.tensor(rng.uniform(2.0, 5.0, size=(NZ, NX)), dtype=torch.float64, requires_grad=True)
    z = torch.tensor(rng.uniform(1.0, 2.0, size=NIF), dtype=torch.float64, requires_grad=True)

    def f(v_, z_):
        return torch.nan_to_num(tt(v_, z_), nan=0.0)

    assert torch.autograd.gradcheck(f, (v, z), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_gradcheck_adjoint_backward_velocity_and_slowness():
    """The adjoint backward's plumbing: the gradient of every grid node and
    of the chosen interface's depths, with the velocity or its reciprocal
    as the input."""
    rng = np.random.default_rng(2)
    z = torch.tensor(rng.uniform(1.0, 2.0, size=NIF), dtype=torch.float64, requires_grad=True)
    for param in ("velocity", "slowness"):
        ctx, tt = make("adjoint", param)
        v = rng.uniform(2.0, 5.0, size=(NZ, NX))
        m = torch.tensor(v if param == "velocity" else 1.0 / v, dtype=torch.float64, requires_grad=True)

        def f(m_, z_):
            return torch.nan_to_num(tt(m_, z_), nan=0.0)

        assert torch.autograd.gradcheck(f, (m, z), eps=1e-7, atol=1e-8, rtol=1e-6)


def test_backward_choices_are_checked():
    import pytest
    with pytest.raises(ValueError):
        make("other")
    with pytest.raises(ValueError):
        make("adjoint", "density")


def test_failed_ray_and_water_nodes_get_no_gradient():
    ctx, tt = make()
    v = torch.full((NZ, NX), 3.0, dtype=torch.float64, requires_grad=True)
    z = torch.zeros(NIF, dtype=torch.float64, requires_grad=True)
    t = tt(v, z)
    assert torch.isnan(t[ctx.bad]) and torch.isfinite(t[:ctx.bad]).all()
    w = torch.zeros(NRAY, dtype=torch.float64)
    w[ctx.bad] = 1.0
    torch.nansum(t * w).backward()
    assert torch.count_nonzero(v.grad) == 0 and torch.count_nonzero(z.grad) == 0
    v.grad = None
    t = tt(v, z)
    t[0].backward()
    assert torch.count_nonzero(v.grad[0]) == 0           # water row
    assert torch.count_nonzero(v.grad[1:]) > 0


def test_backward_after_another_forward_repeats_its_own():
    """A forward at other inputs between a forward and its backward: the
    backward re-runs its own forward first (the stand-in's rays depend on
    nothing, so the context's state is checked instead)."""
    ctx, tt = make()
    v = torch.full((NZ, NX), 3.0, dtype=torch.float64, requires_grad=True)
    z = torch.zeros(NIF, dtype=torch.float64)
    t1 = tt(v, z)
    tt(torch.full((NZ, NX), 4.0, dtype=torch.float64), z)
    assert np.all(ctx.v == 4.0)
    torch.nan_to_num(t1).sum().backward()
    assert np.all(ctx.v == 3.0)
    assert torch.count_nonzero(v.grad[1:]) > 0
