"""The adjoint-state gradient at full size: 2001x301 grid (0.1 km), 60 OBS
every 3.3 km, 1000 shots every 0.2 km, Pg at every offset and PmP at
3-30 km (74 981 rays).  Picks: the receiver times of the +-5 % checkerboard
model (10 km x 3 km cells) plus 30 ms Gaussian noise (seed 4); the
gradient of J = 1/2 sum (t - t_obs)^2 at the starting model (no
checkerboard) with respect to every grid node's slowness and every Moho
node's depth.

* Every reached node of the 180 swept fields is reproduced by its
  recomputed winning update (or head-wave chain), and no partial is dropped.
* Central finite differences of J through the forward and the rays along
  a direction of random signs on 200 random nodes with a gradient (slowness,
  then Moho depth), each within 1e-3 relative, and at 8 single nodes of
  each kind, together within 2e-3 of the gradients' sum of magnitudes.
  Slowness nodes come from below the top two rows (the sea-surface pairs
  of tied updates put kinks in J there).
  The step is h = 1e-6 (s/km, km): at h = 1e-4 the discrete times of a
  0.1 km grid switch between tied updates inside the step, and the
  difference quotient wanders (measured: 2.03 at 1e-4, 0.56 at 1e-5,
  1.22670 at 1e-6 and 1.22740 at 1e-7 against the gradient's 1.22682 for
  the first direction); those quotients are printed beside it.
* Against the reference's ray-based G^T r (compute_frechet_for_ray2d rows,
  src/inversion2d.f90:4548-4787, formed by fmt2d_frechet_apply) over the
  velocity parameters: cosine above 0.95 node by node and above 0.999
  after both are averaged over 1.1 km boxes (measured 0.968 and 0.9997);
  Moho depths: cosine above 0.7 (measured 0.76; the ray rows take the
  reflection-point formula, the adjoint the discrete cut-cell geometry)."""
import os

import numpy as np
import pytest

from conftest import load_case

pytestmark = pytest.mark.gpu

H = 1e-6


@pytest.fixture(scope="module")
def full(tmp_path_factory):
    from fmt2d_amd import formats, hip, synth
    m_true, m0, g = synth.gradient_case()
    d = str(tmp_path_factory.mktemp("adjfull"))
    synth.write_fmt2d_case(d, m_true, g)
    true_in = load_case(d)
    geom = formats.read_geom(os.path.join(d, "geom.dat"))
    synth.write_fmt2d_case(d, m0, g)
    start_in = load_case(d)
    rays = (geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
    print(f"[gpu] adjoint gradient at full size: 2001x301, 60 OBS, {len(geom.rec_x)} rays", flush=True)
    ctx = hip.Context(0)
    obs = list(geom.obs_x)
    ctx.set_model(*true_in)
    ctx.forward(obs)
    t_true, v_true = ctx.trace(*rays)
    t_obs = t_true + np.random.default_rng(4).normal(0.0, 0.03, size=len(t_true))
    ctx.set_model(*start_in)
    ctx.forward(obs)
    t, valid = ctx.trace(*rays)
    ok = valid & v_true
    r = np.where(ok, t - t_obs, 0.0)
    gv, gz, st = ctx.adjoint_gradient(r)
    v0 = ctx.model(0).copy()
    print(f"  adjoint: {st.fields} fields, {st.rounds} rounds, record {st.ms_record:.0f} ms, sweep "
          f"{st.ms_reverse:.0f} ms, total {st.ms_total:.0f} ms, peak {st.peak_bytes / 2**30:.2f} GiB", flush=True)

    def J(v=None, z=None):
        if v is not None:
            ctx.model_put(0, v)
        if z is not None:
            ctx.model_put(4, z)
        ctx.forward(obs)
        tt, _ = ctx.trace(*rays)
        rr = np.where(ok, tt - t_obs, 0.0)
        return 0.5 * float(rr @ rr)

    yield dict(ctx=ctx, gs=-gv * v0 * v0, gz=gz[3], st=st, r=r, J=J, v0=v0, z0=ctx.model(4).copy(), rays=rays,
               obs=obs)
    ctx.close()


def test_every_node_reproduced(full):
    st = full["st"]
    assert st.fields == 180
    assert st.leaves_unmatched == 0
    assert st.dropped_partials == 0


def test_slowness_direction_on_200_nodes(full):
    ctx, gs, J, v0 = full["ctx"], full["gs"], full["J"], full["v0"]
    rng = np.random.default_rng(5)
    pick = rng.choice(np.flatnonzero(gs[2:].ravel() != 0.0) + 2 * v0.shape[1], 200, replace=False)
    sign = rng.choice([-1.0, 1.0], 200)
    d = np.zeros(v0.size)
    d[pick] = sign
    s0 = (1.0 / v0).ravel()
    an = float(gs.ravel()[pick] @ sign)
    try:
        for h in (1e-4, H):
            fd = (J(1.0 / (s0 + h * d).reshape(v0.shape)) - J(1.0 / (s0 - h * d).reshape(v0.shape))) / (2 * h)
            print(f"slowness direction, h={h:g}: FD {fd:.8e}, adjoint {an:.8e}")
    finally:
        ctx.model_put(0, v0)
    assert abs(fd - an) <= 1e-3 * abs(fd)


def test_moho_direction_on_200_nodes(full):
    ctx, g, J, z0 = full["ctx"], full["gz"], full["J"], full["z0"]
    rng = np.random.default_rng(6)
    pick = rng.choice(np.flatnonzero(g != 0.0), 200, replace=False)
    sign = rng.choice([-1.0, 1.0], 200)
    d = np.zeros(len(z0))
    d[pick] = sign
    an = float(g[pick] @ sign)
    try:
        for h in (1e-4, H):
            fd = (J(z=z0 + h * d) - J(z=z0 - h * d)) / (2 * h)
            print(f"Moho direction, h={h:g}: FD {fd:.8e}, adjoint {an:.8e}")
    finally:
        ctx.model_put(4, z0)
    assert abs(fd - an) <= 1e-3 * abs(fd)


def test_single_nodes(full):
    ctx, gs, g, J, v0, z0 = (full[k] for k in ("ctx", "gs", "gz", "J", "v0", "z0"))
    rng = np.random.default_rng(7)
    err, mag = [], []
    try:
        big = np.abs(gs) > 1e-2 * np.abs(gs).max()
        big[:2] = False
        for k in rng.choice(np.flatnonzero(big.ravel()), 8, replace=False):
            j, i = divmod(int(k), v0.shape[1])
            s = 1.0 / v0[j, i]
            vp, vm = v0.copy(), v0.copy()
            vp[j, i] = 1.0 / (s + H)
            vm[j, i] = 1.0 / (s - H)
            fd = (J(vp) - J(vm)) / (2 * H)
            print(f"node ({i},{j}): FD {fd:.6e}, adjoint {gs[j, i]:.6e}")
            err.append(abs(fd - gs[j, i]))
            mag.append(abs(gs[j, i]))
        ctx.model_put(0, v0)
        for c in rng.choice(np.flatnonzero(np.abs(g) > 1e-2 * np.abs(g).max()), 8, replace=False):
            zp, zm = z0.copy(), z0.copy()
            zp[c] += H
            zm[c] -= H
            fd = (J(z=zp) - J(z=zm)) / (2 * H)
            print(f"Moho node {c}: FD {fd:.6e}, adjoint {g[c]:.6e}")
            err.append(abs(fd - g[c]))
            mag.append(abs(g[c]))
    finally:
        ctx.model_put(0, v0)
        ctx.model_put(4, z0)
    assert sum(err) <= 2e-3 * sum(mag)


def test_against_ray_based_gradient(full):
    from scipy.ndimage import uniform_filter
    ctx, gs, g, r = full["ctx"], full["gs"], full["gz"], full["r"]
    ctx.set_inversion(True, [4])
    ctx.forward(full["obs"])
    ctx.trace(*full["rays"])
    y = ctx.frechet_apply(r, transpose=True)
    idx = np.asarray(ctx.param_index())
    n_vel, _ = ctx.inversion_info()
    sel = (idx > 0) & (idx <= n_vel)
    gr = np.zeros(gs.shape)
    gr[sel] = y[idx[sel] - 1]

    def cos(a, b):
        return float(np.sum(a * b) / np.linalg.norm(a) / np.linalg.norm(b))

    c_node = cos(gs[sel], gr[sel])
    c_box = cos(uniform_filter(np.where(sel, gs, 0.0), 11), uniform_filter(gr, 11))
    c_moho = cos(g, y[n_vel:])
    print(f"cosine against G^T r: velocity {c_node:.4f} node by node, {c_box:.5f} over 1.1 km boxes; "
          f"Moho {c_moho:.3f}")
    assert c_node > 0.95 and c_box > 0.999 and c_moho > 0.7
