"""The adjoint-state gradient (fmt2d_adjoint_gradient) on the 1 km
checkerboard case with all five phases (3 OBS, 1376 rays): the misfit
J = 1/2 sum (t - t_true)^2 of the starting model's receiver times against
the checkerboard model's.

* Every reached node of every swept field is reproduced by its recomputed
  winning update (or head-wave chain), bit for bit, and no partial
  derivative is dropped.
* Against central finite differences of J through the forward and the rays
  (h = 1e-6 in slowness, 1e-6 km in depth; the forward is deterministic):
  12 velocity nodes below the top two rows, 6 Moho nodes and 6
  sediment-base nodes, each within 1e-4 relative plus 1e-6 of the largest
  gradient entry of its kind (measured: at most 1.1e-4 relative, at a Moho
  node whose gradient is 1/150 of the largest).  The top two rows are left
  out: the sea-surface row's grid nodes coincide with the surface's
  interface nodes, which take the velocities of the row below, and each
  pair keeps the smaller of two tied updates, so J has kinks there
  (measured at (73,1): one-sided slopes -0.686 and -0.944 at h = 1e-6 and
  1e-8 alike; the gradient, -0.829, lies between them).
* Batches: a memory cap that sweeps one field at a time gives the same
  gradient (sums in another order: 1e-12 relative).
* torch.autograd: TravelTimes(backward="adjoint") returns the same gradient
  with the velocity as input, and -v^2 times it with the slowness."""
import os

import numpy as np
import pytest

from conftest import load_case

pytestmark = pytest.mark.gpu

H_S, H_Z = 1e-6, 1e-6


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    from fmt2d_amd import formats, hip, synth
    m_true, m0, g = synth.inversion_case(dx=1.0, phases=True)
    d = str(tmp_path_factory.mktemp("adj"))
    synth.write_fmt2d_case(d, m_true, g)
    true_in = load_case(d)
    geom = formats.read_geom(os.path.join(d, "geom.dat"))
    synth.write_fmt2d_case(d, m0, g)
    start_in = load_case(d)
    rays = (geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
    print("[gpu] adjoint gradient: 1 km checkerboard case, five phases", flush=True)
    ctx = hip.Context(0)
    ctx.set_model(*true_in)
    ctx.forward(list(geom.obs_x))
    t_true, v_true = ctx.trace(*rays)
    ctx.set_model(*start_in)
    obs = list(geom.obs_x)
    ctx.forward(obs)
    t, valid = ctx.trace(*rays)
    ok = valid & v_true
    r = np.where(ok, t - t_true, 0.0)
    gv, gz, st = ctx.adjoint_gradient(r)
    print(f"  {st.fields} fields, {st.rounds} rounds, {st.ms_total:.1f} ms", flush=True)

    def J(v=None, z=None, k=4):
        if v is not None:
            ctx.model_put(0, v)
        if z is not None:
            ctx.model_put(k, z)
        ctx.forward(obs)
        tt, _ = ctx.trace(*rays)
        rr = np.where(ok, tt - t_true, 0.0)
        return 0.5 * float(rr @ rr)

    yield dict(ctx=ctx, gv=gv, gz=gz, st=st, r=r, J=J, v0=ctx.model(0).copy(), rays=rays, obs=obs, ok=ok,
               t_true=t_true, t=t)
    ctx.close()


def test_every_node_reproduced(case):
    st = case["st"]
    assert st.fields > 0 and st.rounds > 0
    assert st.leaves_unmatched == 0
    assert st.dropped_partials == 0


def test_velocity_nodes_against_finite_differences(case):
    ctx, gv, J, v0 = case["ctx"], case["gv"], case["J"], case["v0"]
    gs = -gv * v0 * v0
    big = np.abs(gs) > 1e-3 * np.abs(gs).max()
    big[:2, :] = False
    rng = np.random.default_rng(5)
    worst = 0.0
    try:
        for k in rng.choice(np.flatnonzero(big.ravel()), 12, replace=False):
            j, i = divmod(int(k), v0.shape[1])
            s = 1.0 / v0[j, i]
            vp, vm = v0.copy(), v0.copy()
            vp[j, i] = 1.0 / (s + H_S)
            vm[j, i] = 1.0 / (s - H_S)
            fd = (J(vp) - J(vm)) / (2 * H_S)
            rel = abs(fd - gs[j, i]) / abs(fd)
            print(f"node ({i},{j}) v={v0[j, i]:.3f}: FD {fd:.8e} adjoint {gs[j, i]:.8e} rel {rel:.1e}")
            worst = max(worst, abs(fd - gs[j, i]) - 1e-4 * abs(fd) - 1e-6 * np.abs(gs).max())
    finally:
        ctx.model_put(0, v0)
    assert worst <= 0.0


@pytest.mark.parametrize("k", [4, 3])
def test_interface_nodes_against_finite_differences(case, k):
    ctx, gz, J = case["ctx"], case["gz"], case["J"]
    z0 = ctx.model(k).copy()
    g = gz[k - 1]
    rng = np.random.default_rng(6 + k)
    worst = 0.0
    try:
        for c in rng.choice(np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max()), 6, replace=False):
            zp, zm = z0.copy(), z0.copy()
            zp[c] += H_Z
            zm[c] -= H_Z
            fd = (J(z=zp, k=k) - J(z=zm, k=k)) / (2 * H_Z)
            rel = abs(fd - g[c]) / abs(fd)
            print(f"interface {k} node {c}: FD {fd:.8e} adjoint {g[c]:.8e} rel {rel:.1e}")
            worst = max(worst, abs(fd - g[c]) - 1e-4 * abs(fd) - 1e-6 * np.abs(g).max())
    finally:
        ctx.model_put(k, z0)
    assert worst <= 0.0


def test_batched_sweep_gives_the_same_gradient(case):
    ctx, rays, obs = case["ctx"], case["rays"], case["obs"]
    ctx.forward(obs)
    ctx.trace(*rays)
    nslot = (ctx.info.nz + 2) * ctx.info.nx
    try:
        ctx.set_adjoint_memory(nslot * 120)
        gv, gz, st = ctx.adjoint_gradient(case["r"])
    finally:
        ctx.set_adjoint_memory(0)
    assert st.leaves_unmatched == 0
    for a, b in ((gv, case["gv"]), (gz, case["gz"])):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())


def test_autograd_adjoint_backward(case):
    """The backward against a direct call at the velocities the forward ran
    with (1 / (1 / v) is not always v: a last-bit change can flip a tie at
    the sea surface)."""
    import torch
    from fmt2d_amd.autograd import TravelTimes
    ctx, rays, ok, obs = case["ctx"], case["rays"], case["ok"], case["obs"]
    t_true = torch.tensor(case["t_true"], dtype=torch.float64, device="cuda:0")
    mask = torch.tensor(ok, device="cuda:0")
    v0 = case["v0"]
    z3_0, z4_0 = ctx.model(3).copy(), ctx.model(4).copy()
    try:
        for param in ("velocity", "slowness"):
            tt = TravelTimes(ctx, obs, *rays, invert_velocity=True, ifaces=[3, 4], param=param)
            m_np = v0 if param == "velocity" else 1.0 / v0
            m = torch.tensor(m_np, dtype=torch.float64, device="cuda:0", requires_grad=True)
            z3 = torch.tensor(z3_0, dtype=torch.float64, device="cuda:0", requires_grad=True)
            z4 = torch.tensor(z4_0, dtype=torch.float64, device="cuda:0", requires_grad=True)
            t = tt(m, z3, z4)
            r = torch.where(mask, t - t_true, torch.zeros_like(t))
            (0.5 * torch.sum(r * r)).backward()
            assert tt.last_stats.leaves_unmatched == 0
            v_used = tt.velocity_of(m_np)
            ctx.model_put(0, v_used)
            ctx.forward(obs)
            tr, _ = ctx.trace(*rays)
            gv, gz, _ = ctx.adjoint_gradient(np.where(ok, tr - case["t_true"], 0.0))
            want = gv if param == "velocity" else -gv * v_used * v_used
            got = m.grad.cpu().numpy()
            assert np.allclose(got, want, rtol=1e-12, atol=1e-14 * np.abs(want).max()), \
                f"{param}: max |diff| {np.abs(got - want).max():.3e} of {np.abs(want).max():.3e}"
            for z, k in ((z3, 3), (z4, 4)):
                assert np.allclose(z.grad.cpu().numpy(), gz[k - 1], rtol=1e-12, atol=1e-14 * np.abs(gz[k - 1]).max())
    finally:
        ctx.model_put(0, v0)
