"""fmt2d_amd.autograd with the ray-based backward (backward="rays") on the
GPU (all five phases on the 1 km checkerboard model, velocity grid +
sediment base + Moho as parameters).

* The adjoint: fmt2d_frechet_apply (G^T x and G x on the device, each sum
  in stored order) equals sequential sums over the rows that
  fmt2d_frechet_csr returns, bit for bit; the backward of a weighted sum of
  the receiver times equals -(G^T w)/v^2 on the velocity nodes and G^T w on
  the interface nodes.
* Against finite differences of the eikonal times: the directional
  derivative along a smooth velocity perturbation (central differences of
  the forward) agrees with the gradient's within 5 %: the ray-theoretic
  linearisation the reference inverts with, against the solver's own times.
  The interface rows assume the velocities move with the interface, while
  the grid's node velocities stay where they are when only the depths
  change; the interface check therefore uses a model with one velocity
  below the seafloor, where moving the Moho changes only the reflection
  geometry, and sums the PmP rays of the OBS above the Moho perturbation
  (within 5 %).  Left out: the sediment base (its 1-2 km thick layer spans
  one or two grid rows at 1 km spacing, so its discrete reflection times
  do not follow the ray geometry) and the PmP rays of the other OBS (one
  of them switches branch, a jump of 0.4 s, under a 0.1 m depth change)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(tmp_path_factory):
    import torch
    from fmt2d_amd import formats, hip, synth
    from fmt2d_amd.autograd import TravelTimes
    d = str(tmp_path_factory.mktemp("ag"))
    m_true, m0, g = synth.inversion_case(dx=1.0, phases=True)
    synth.write_fmt2d_case(d, m_true, g)
    sm = formats.read_smesh(os.path.join(d, "model.smesh"))
    sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
    mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
    geom = formats.read_geom(os.path.join(d, "geom.dat"))
    print("[gpu] autograd: all phases, 1 km grid, velocity + sediment base + Moho", flush=True)
    ctx = hip.Context(0)
    ctx.set_model(sm, sx, sz, mx, mz)
    tt = TravelTimes(ctx, geom.obs_x, geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z, True, [3, 4],
                     backward="rays")
    dev = "cuda:0"
    v = torch.tensor(ctx.model(0), dtype=torch.float64, device=dev)
    z3 = torch.tensor(ctx.model(3), dtype=torch.float64, device=dev)
    z4 = torch.tensor(ctx.model(4), dtype=torch.float64, device=dev)
    yield dict(ctx=ctx, tt=tt, v=v, z3=z3, z4=z4, nray=len(geom.rec_x), path=np.asarray(geom.ray_type),
               obs=np.asarray(geom.obs_id))
    ctx.close()


def test_adjoint_sums_in_order(setup):
    ctx, tt, v, z3, z4 = (setup[k] for k in ("ctx", "tt", "v", "z3", "z4"))
    t = tt(v, z3, z4)
    rp, col, val = ctx.frechet_csr(index_base=0)
    nray, n_par = len(rp) - 1, tt.n_par
    rng = np.random.default_rng(5)
    x = rng.normal(size=nray)
    y = ctx.frechet_apply(x, transpose=True)
    G = sp.csr_matrix((val, col, rp), shape=(nray, n_par))
    Gc = G.tocsc()
    Gc.sort_indices()
    for j in rng.choice(n_par, 200, replace=False):
        a, b = Gc.indptr[j], Gc.indptr[j + 1]
        s = 0.0
        for k in range(a, b):
            s = s + Gc.data[k] * x[Gc.indices[k]]
        assert y[j] == s, f"column {j}"
    m = rng.normal(size=n_par)
    yr = ctx.frechet_apply(m, transpose=False)
    for i in rng.choice(nray, 200, replace=False):
        s = 0.0
        for k in range(rp[i], rp[i + 1]):
            s = s + val[k] * m[col[k]]
        assert yr[i] == s, f"row {i}"
    # failed rays have NaN times and empty rows
    empty = np.diff(rp) == 0
    assert np.all(empty[~np.isfinite(t.cpu().numpy())])


def test_backward_is_the_adjoint(setup):
    import torch
    ctx, tt = setup["ctx"], setup["tt"]
    v = setup["v"].clone().requires_grad_(True)
    z3 = setup["z3"].clone().requires_grad_(True)
    z4 = setup["z4"].clone().requires_grad_(True)
    t = tt(v, z3, z4)
    w = torch.tensor(np.random.default_rng(7).normal(size=t.numel()), dtype=torch.float64, device=t.device)
    torch.nansum(w * t).backward()
    valid = torch.isfinite(t).cpu().numpy()
    rp, col, val = ctx.frechet_csr(index_base=0)
    G = sp.csr_matrix((val, col, rp), shape=(len(rp) - 1, tt.n_par))
    y = G.T @ np.where(valid, w.cpu().numpy(), 0.0)
    idx = tt.node_param
    sel = (idx > 0) & (idx <= tt.n_vel)
    vv = v.detach().cpu().numpy()
    gv = np.zeros(vv.shape)
    gv[sel] = -y[idx[sel] - 1] / vv[sel] ** 2
    got = v.grad.cpu().numpy()
    assert np.allclose(got, gv, rtol=1e-12, atol=1e-14 * np.abs(gv).max())
    assert np.count_nonzero(got[~sel]) == 0
    for z, sl in ((z3, tt.iface_slices[0]), (z4, tt.iface_slices[1])):
        assert np.allclose(z.grad.cpu().numpy(), y[sl], rtol=1e-12, atol=1e-14 * max(np.abs(y[sl]).max(), 1e-30))


def directional(setup, dv=None, dz3=None, dz4=None, eps=1.0, v_base=None, phase=None, obs=None):
    import torch
    tt = setup["tt"]
    v = (setup["v"] if v_base is None else v_base).clone().requires_grad_(True)
    z3 = setup["z3"].clone().requires_grad_(True)
    z4 = setup["z4"].clone().requires_grad_(True)
    w_np = np.random.default_rng(11).uniform(0.5, 1.5, size=setup["nray"])
    if phase is not None:
        w_np[setup["path"] != phase] = 0.0
    if obs is not None:
        w_np[setup["obs"] != obs] = 0.0
    w = torch.tensor(w_np, dtype=torch.float64, device=v.device)
    t0 = tt(v, z3, z4)
    def d(a, like):
        return torch.zeros_like(like) if a is None else torch.tensor(a, dtype=torch.float64, device=v.device)
    dvt, d3, d4 = d(dv, v), d(dz3, z3), d(dz4, z4)
    with torch.no_grad():
        tp = tt(v + eps * dvt, z3 + eps * d3, z4 + eps * d4)
        tm = tt(v - eps * dvt, z3 - eps * d3, z4 - eps * d4)
    ok = torch.isfinite(t0) & torch.isfinite(tp) & torch.isfinite(tm)
    torch.sum(torch.where(ok, w * t0, torch.zeros_like(t0))).backward()
    fd = float(torch.sum(torch.where(ok, w * (tp - tm), torch.zeros_like(tp))) / (2 * eps))
    an = float(torch.sum(v.grad * dvt) + torch.sum(z3.grad * d3) + torch.sum(z4.grad * d4))
    return fd, an, int(ok.sum())


def test_velocity_gradient_against_finite_differences(setup):
    tt = setup["tt"]
    info = setup["ctx"].info
    x = info.x0 + info.dx * np.arange(info.nx)
    z = info.z0 + info.dz * np.arange(info.nz)
    X, Z = np.meshgrid(x, z)
    bump = np.exp(-((X - 95.0) / 25.0) ** 2 - ((Z - 10.0) / 4.0) ** 2)
    bump[tt.node_param == 0] = 0.0
    fd, an, n = directional(setup, dv=0.02 * bump, eps=1.0)
    print(f"velocity bump: finite differences {fd:.6e}, gradient {an:.6e} over {n} rays", flush=True)
    assert abs(fd - an) <= 0.05 * abs(fd)


def test_moho_gradient_against_finite_differences(setup):
    import torch
    tt = setup["tt"]
    v_one = setup["v"].clone()
    v_one[torch.from_numpy(tt.node_param > 0).to(v_one.device)] = 6.0
    x = np.arange(len(setup["z4"].cpu().numpy())) * setup["ctx"].info.dx + setup["ctx"].info.x0
    dz = 0.05 * np.exp(-((x - 95.0) / 20.0) ** 2)
    fd, an, n = directional(setup, eps=1.0, v_base=v_one, dz4=dz, phase=3, obs=2)
    print(f"Moho bump, PmP of the OBS at 95 km: finite differences {fd:.6e}, gradient {an:.6e}", flush=True)
    assert fd > 0.0
    assert abs(fd - an) <= 0.05 * abs(fd)
