"""The joint velocity + interface inversion at full size: 2001x301 grid
(0.1 km), 200 OBS every 1 km, shots every 0.2 km, five phases in the
reference's windows (434 k rays), picks from the +-5 % checkerboard model
with 50 ms noise (seed 5); velocity parameters on a 0.5 km mesh (401 x 61
mesh nodes, every 5th grid node), sediment-base and Moho depths every 1 km
(2 x 201 nodes, every 10th interface node); hybrid smoothing, damping 0.1.

The same properties as the coarse mesh test (tests/test_mesh_gpu.py), at
this size:

* the mesh and parameter counts;
* the system's gtg_diag equals the column sums of squares of G P, with G
  the grid's Frechet rows (fmt2d_frechet_csr) and P built here from the node
  coordinates (bilinear hats of the velocity mesh, linear hats along the
  interfaces), columns scaled as the system scales them;
* one iteration (fmt2d_damped_update, fmt2d_line_search) lowers the RMS and
  moves the grid's slowness only within the range of P."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

SV, SI = 5, 10
IFACES = [3, 4]
REFL_WEIGHT = 10.0


def hat(u):
    return np.maximum(0.0, 1.0 - np.abs(u))


@pytest.fixture(scope="module")
def joint(tmp_path_factory):
    import os
    from fmt2d_amd import formats, hip, synth
    obs_x = np.round(np.arange(0.5, 200.0, 1.0), 6)
    m_true, m0, g = synth.inversion_case(dx=0.1, obs_x=obs_x, shot_dx=0.2, phases=True)
    d = str(tmp_path_factory.mktemp("joint"))

    def load(m):
        synth.write_fmt2d_case(d, m, g)
        return (formats.read_smesh(os.path.join(d, "model.smesh")), *formats.read_refl(os.path.join(d, "sed.dat")),
                *formats.read_refl(os.path.join(d, "moho.dat")))
    true_in = load(m_true)
    geom = formats.read_geom(os.path.join(d, "geom.dat"))
    start_in = load(m0)
    rays = (geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
    print(f"[gpu] joint inversion at full size: {len(geom.rec_x)} rays, mesh strides {SV}/{SI}", flush=True)
    ctx = hip.Context(0)
    ctx.set_model(*true_in)
    ctx.forward(list(geom.obs_x))
    t_true, v_true = ctx.trace(*rays)
    noise = np.random.default_rng(5).normal(0.0, 0.05, size=len(t_true))
    t_obs = np.where(v_true, t_true + noise, -1.0)
    ctx.set_model(*start_in)
    ctx.set_inversion(True, IFACES)
    ctx.set_inversion_mesh(SV, SI)
    ctx.forward(list(geom.obs_x))
    t, valid = ctx.trace(*rays)
    nd, nnz, _ = ctx.build_system(t_obs, crit_chi=1.0, refl_weight=REFL_WEIGHT)
    print(f"  system: {nd} rows, {nnz} non-zeros", flush=True)
    yield dict(ctx=ctx, obs_x=list(geom.obs_x), rays=rays, t_obs=t_obs, t=t, valid=valid)
    ctx.close()


def mesh_p(ctx):
    """P from node coordinates (vectorised): the mesh's velocity parameters
    are the grid's parameter nodes at every SV-th column and row; every grid
    parameter node takes the bilinear hats of the (up to four) mesh
    parameters around it; interface node ix takes the linear hats of the
    coarse nodes ix // SI and ix // SI + 1."""
    idx = np.asarray(ctx.param_index())
    nz, nx = idx.shape
    mesh = np.full((nz, nx), -1, dtype=np.int64)
    mz, mx = np.meshgrid(np.arange(0, nz, SV), np.arange(0, nx, SV), indexing="ij")
    is_par = idx[mz, mx] > 0
    mesh[mz[is_par], mx[is_par]] = np.arange(int(is_par.sum()))
    n_vel_m = int(is_par.sum())
    jz, jx = np.nonzero(idx > 0)
    rows_f = idx[jz, jx] - 1
    rows, cols, vals = [], [], []
    for cz in (0, 1):
        for cx in (0, 1):
            kz = (jz // SV + cz) * SV
            kx = (jx // SV + cx) * SV
            ok = (kz < nz) & (kx < nx)
            kz, kx, r, z0, x0 = kz[ok], kx[ok], rows_f[ok], jz[ok], jx[ok]
            col = mesh[kz, kx]
            w = hat((x0 - kx) / SV) * hat((z0 - kz) / SV)
            keep = (col >= 0) & (w > 0.0)
            rows.append(r[keep]); cols.append(col[keep]); vals.append(w[keep])
    n_vel_f = int(idx.max())
    jf, km = n_vel_f, n_vel_m
    for k_if in IFACES:
        nn = len(ctx.model(k_if))
        nc = (nn - 1) // SI + 1
        ix = np.arange(nn)
        for c in (ix // SI, ix // SI + 1):
            w = hat((ix - c * SI) / SI)
            keep = (c < nc) & (w > 0.0)
            rows.append(jf + ix[keep]); cols.append(km + c[keep]); vals.append(w[keep])
        jf += nn
        km += nc
    P = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(jf, km))
    return P, mesh, n_vel_m


def test_counts(joint):
    ctx = joint["ctx"]
    vnx, vnz, n_fre = ctx.mesh_info()
    assert (vnx, vnz) == (401, 61)
    P, mesh, n_vel_m = mesh_p(ctx)
    assert ctx.inversion_info() == (n_vel_m, P.shape[1])
    assert n_fre == P.shape[0]
    assert P.shape[1] == n_vel_m + 2 * 201


def test_system_is_g_times_p(joint):
    ctx = joint["ctx"]
    P, mesh, n_vel_m = mesh_p(ctx)
    rp, col, val = ctx.frechet_csr(index_base=0)
    G = sp.csr_matrix((val, col, rp), shape=(len(rp) - 1, P.shape[0]))
    t, valid, t_obs = joint["t"], joint["valid"], joint["t_obs"]
    keep = valid.astype(bool) & (t_obs >= 0.0) & (np.abs(t_obs - t) <= 1.0)
    v_init = ctx.model(0)
    mvs = np.full(P.shape[1], REFL_WEIGHT)
    mz, mx = np.nonzero(mesh >= 0)
    mvs[mesh[mz, mx]] = 1.0 / np.maximum(v_init[mz, mx], 1.0)
    GP = (G[np.flatnonzero(keep)] @ P) @ sp.diags(mvs)
    want = np.asarray(GP.multiply(GP).sum(axis=0)).ravel()
    gtg, res = ctx.system()
    assert len(res) == int(keep.sum())
    assert np.allclose(gtg, want, rtol=1e-11, atol=1e-13 * want.max())
    # the 5000-point section cap keeps the valid rays to the upper part of
    # the model: 46 % of the mesh velocity parameters lie under rays
    assert np.count_nonzero(want[:n_vel_m]) > 0.3 * n_vel_m


def test_iteration_lowers_rms_in_the_range_of_p(joint):
    ctx = joint["ctx"]
    P, mesh, n_vel_m = mesh_p(ctx)
    v0 = ctx.model(0).copy()
    dm, d = ctx.damped_update(lam_damp=0.1, lam_damp_iface=0.1, smooth_type=3, w_smooth=0.3, w_smooth_iface=0.3,
                              max_iter=400, corr_x=10.0, corr_z=3.0, corr_x_iface=8.0)
    vmask = np.ascontiguousarray(joint["valid"], dtype=np.int32)
    dm_step, r = ctx.line_search(joint["t_obs"], dm, crit_chi=1.0, valid=vmask)
    print(f"full-size joint iteration: {d.iterations} LSQR iterations, rms {r.rms_orig:.5f} -> {r.rms_new:.5f}, "
          f"step {r.step}, {r.forwards} trial forwards", flush=True)
    assert r.step > 0.0 and r.rms_new < r.rms_orig
    v1 = ctx.model(0)
    idx = np.asarray(ctx.param_index())
    jz, jx = np.nonzero(idx > 0)
    ds = 1.0 / v1[jz, jx] - 1.0 / v0[jz, jx]
    mz, mx = np.nonzero(mesh >= 0)
    ds_mesh = np.zeros(n_vel_m)
    ds_mesh[mesh[mz, mx]] = 1.0 / v1[mz, mx] - 1.0 / v0[mz, mx]
    n_vel_f = int(idx.max())
    pred = P[:n_vel_f, :n_vel_m] @ ds_mesh
    got = np.zeros(n_vel_f)
    got[idx[jz, jx] - 1] = ds
    vv0, vv1 = np.zeros(n_vel_f), np.zeros(n_vel_f)
    vv0[idx[jz, jx] - 1] = v0[jz, jx]
    vv1[idx[jz, jx] - 1] = v1[jz, jx]
    # nodes where a clamp acts are left out (below 2 km/s, 30 % from the
    # start), and so are the nodes whose prolongation reads a clamped mesh node
    def unclamped(a, b):
        return (a >= 2.0) & (0.7 * a < b) & (b < 1.3 * a) & (2.0 < b) & (b < 10.0)
    ok = unclamped(vv0, vv1)
    mesh_ok = np.zeros(n_vel_m, dtype=bool)
    mesh_ok[mesh[mz, mx]] = unclamped(v0[mz, mx], v1[mz, mx])
    ok &= (P[:n_vel_f, :n_vel_m] @ (~mesh_ok).astype(float)) == 0.0
    assert ok.sum() > 0.9 * n_vel_f
    assert np.abs(got - pred)[ok].max() <= 1e-12 * max(np.abs(got).max(), 1e-30)
    assert np.abs(got).max() > 0.0
