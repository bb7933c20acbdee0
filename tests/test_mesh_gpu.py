"""The inversion mesh coarser than the propagation grid
(fmt2d_set_inversion_mesh) on the 1 km all-phase case (201 x 31 nodes):
velocity parameters every 5 km (41 x 7 mesh nodes), sediment-base and Moho
depth parameters every 10 km (21 nodes each).  The reference has no such
parameterisation (its inversion grid is the propagation grid); at strides
1 the mesh is the grid, and every byte-identical inversion test runs
through the same G P code with P = I.  Here:

* a model step of one mesh parameter moves the grid's slowness by that
  parameter's hat function (1 at its node, falling linearly to 0 at the
  neighbouring mesh nodes in x and in z, the product of the two), and an
  interface parameter moves the interface's grid-column nodes by the
  one-dimensional hat: what bilinear and linear prolongation are;
* the system's gtg_diag equals the column sums of squares of G P, with G
  the grid's Frechet rows (fmt2d_frechet_csr) and P built here from the
  node coordinates and hat functions, columns scaled as the system scales
  them;
* one iteration (automatic damping, line search) moves the grid's
  slowness only within the range of P, and lowers the RMS; the next
  forward reuses the accepted trial's fields and traces the same rays, bit
  for bit, as a full solve;
* invert2d_hip with the mesh line in invert2d.in runs the alternating
  joint inversion and lowers the RMS."""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fmt2d_amd", "fortran", "invert2d_hip")
SV, SI = 5, 10
IFACES = [3, 4]
REFL_WEIGHT = 10.0


def hat(u):
    return np.maximum(0.0, 1.0 - np.abs(u))


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    from fmt2d_amd import formats, hip, synth
    m_true, m0, g = synth.inversion_case(dx=1.0, phases=True)
    dt = str(tmp_path_factory.mktemp("true"))
    d0 = str(tmp_path_factory.mktemp("start"))
    synth.write_fmt2d_case(dt, m_true, g)
    synth.write_fmt2d_case(d0, m0, g)

    def load(d):
        return (formats.read_smesh(os.path.join(d, "model.smesh")), *formats.read_refl(os.path.join(d, "sed.dat")),
                *formats.read_refl(os.path.join(d, "moho.dat")))
    geom = formats.read_geom(os.path.join(d0, "geom.dat"))
    rays = (geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
    print(f"[gpu] inversion mesh: strides {SV} (velocity), {SI} (interfaces), 1 km all-phase case", flush=True)
    ctx = hip.Context(0)
    ctx.set_model(*load(dt))
    ctx.forward(geom.obs_x)
    t_true, v_true = ctx.trace(*rays)
    ctx.set_model(*load(d0))
    ctx.set_inversion(True, IFACES)
    ctx.set_inversion_mesh(SV, SI)
    yield dict(ctx=ctx, obs_x=geom.obs_x, rays=rays, t_obs=np.where(v_true, t_true, -1.0), dir0=d0)
    ctx.close()


def forward_and_system(c, crit_chi=100.0):
    ctx = c["ctx"]
    ctx.forward(c["obs_x"])
    t, valid = ctx.trace(*c["rays"])
    ctx.build_system(c["t_obs"], crit_chi=crit_chi, refl_weight=REFL_WEIGHT)
    return t, valid


def mesh_p(ctx):
    """P from coordinates: hat functions of the mesh nodes that are
    parameters, on the grid's parameter nodes and interface columns."""
    info = ctx.info
    idx = np.asarray(ctx.param_index())
    nz, nx = idx.shape
    x = info.x0 + info.dx * np.arange(nx)
    z = info.dz * np.arange(nz)
    mesh_nodes = []          # (ix, iz) of the mesh's velocity parameters, numbering order
    for iz in range(0, nz, SV):
        for ix in range(0, nx, SV):
            if idx[iz, ix] > 0:
                mesh_nodes.append((ix, iz))
    n_vel_f = int(idx.max())
    H = SV * info.dx
    rows, cols, vals = [], [], []
    fine = [(ix, iz) for iz in range(nz) for ix in range(nx) if idx[iz, ix] > 0]
    for j, (ix, iz) in enumerate(fine):
        assert idx[iz, ix] == j + 1
        for k, (kx, kz) in enumerate(mesh_nodes):
            w = hat((x[ix] - x[kx]) / H) * hat((z[iz] - z[kz]) / (SV * info.dz))
            if w > 0.0:
                rows.append(j); cols.append(k); vals.append(w)
    n_vel_m = len(mesh_nodes)
    jf, km = n_vel_f, n_vel_m
    for k_if in IFACES:
        nn = len(ctx.model(k_if))
        nc = (nn - 1) // SI + 1
        for ix in range(nn):
            for c in range(nc):
                w = hat((ix - c * SI) / SI)
                if w > 0.0:
                    rows.append(jf + ix); cols.append(km + c); vals.append(w)
        jf += nn
        km += nc
    P = sp.csr_matrix((vals, (rows, cols)), shape=(jf, km))
    return P, mesh_nodes, n_vel_m


def test_counts(case):
    ctx = case["ctx"]
    vnx, vnz, n_fre = ctx.mesh_info()
    assert (vnx, vnz) == (41, 7)
    P, nodes, n_vel_m = mesh_p(ctx)
    assert ctx.inversion_info() == (n_vel_m, P.shape[1])
    assert n_fre == P.shape[0]
    assert P.shape[1] == n_vel_m + 2 * 21


def test_model_step_is_the_hat_function(case):
    ctx = case["ctx"]
    forward_and_system(case)
    P, nodes, n_vel_m = mesh_p(ctx)
    n_par = P.shape[1]
    info = ctx.info
    v0 = ctx.model(0).copy()
    z4 = ctx.model(4).copy()
    k = nodes.index((100, 20))                  # x 100 km, 20 km below the grid top
    dm = np.zeros(n_par)
    dm[k] = 0.05
    ctx.apply_model_step(1.0, 0.0, dm)
    v1 = ctx.model(0)
    mvs = 1.0 / max(v0[20, 100], 1.0)
    nz, nx = v0.shape
    X = np.arange(nx)[None, :]
    Z = np.arange(nz)[:, None]
    want = hat((X - 100) / SV) * hat((Z - 20) / SV) * (1.0 * mvs * 0.05)
    idx = np.asarray(ctx.param_index())
    ok = (idx > 0) & (v0 >= 2.0)                 # the step clamps velocities below 2 km/s to 2
    got = 1.0 / v1 - 1.0 / v0
    assert np.abs(got[ok] - want[ok]).max() <= 1e-14
    assert np.count_nonzero(want[ok]) == 9 * 9
    ctx.model_put(0, v0)
    # Moho parameter c = 9 (x = 90 km): the interface's columns 80..100 km
    c = 9
    dm = np.zeros(n_par)
    dm[n_vel_m + 21 + c] = 0.001
    ctx.apply_model_step(0.0, 1.0, dm)
    dz = ctx.model(4) - z4
    want = hat((np.arange(len(z4)) - c * SI) / SI) * (REFL_WEIGHT * 0.001)
    assert np.abs(dz - want).max() <= 1e-14
    assert np.count_nonzero(want) == 2 * SI - 1
    ctx.model_put(4, z4)


def test_system_is_g_times_p(case):
    ctx = case["ctx"]
    t, valid = forward_and_system(case)
    P, nodes, n_vel_m = mesh_p(ctx)
    rp, col, val = ctx.frechet_csr(index_base=0)
    G = sp.csr_matrix((val, col, rp), shape=(len(rp) - 1, P.shape[0]))
    t_obs = case["t_obs"]
    keep = valid.astype(bool) & (t_obs >= 0.0) & (np.abs(t_obs - t) <= 100.0)
    v_init = ctx.model(0)
    mvs = np.full(P.shape[1], REFL_WEIGHT)
    for k, (ix, iz) in enumerate(nodes):
        mvs[k] = 1.0 / max(v_init[iz, ix], 1.0)
    GP = (G[np.flatnonzero(keep)] @ P) @ sp.diags(mvs)
    want = np.asarray(GP.multiply(GP).sum(axis=0)).ravel()
    gtg, res = ctx.system()
    assert len(res) == int(keep.sum())
    assert np.allclose(gtg, want, rtol=1e-11, atol=1e-13 * want.max())
    # every mesh velocity parameter under the rays carries some weight
    assert np.count_nonzero(want[:n_vel_m]) > 0.5 * n_vel_m


def test_chunked_assembly_is_identical(case):
    """Frechet rows and G P assembled in chunks of a few rays give the same
    arrays, bit for bit, as one assembly of all rays."""
    ctx = case["ctx"]
    forward_and_system(case)
    one = ctx.frechet_csr(index_base=0)
    g1, r1 = ctx.system()
    ctx.set_frechet_chunk(20000)
    try:
        ctx.build_system(case["t_obs"], crit_chi=100.0, refl_weight=REFL_WEIGHT)
        many = ctx.frechet_csr(index_base=0)
        g2, r2 = ctx.system()
    finally:
        ctx.set_frechet_chunk()
    for a, b in zip(one, many):
        assert a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert np.array_equal(g1.view(np.uint64), g2.view(np.uint64))
    assert np.array_equal(r1.view(np.uint64), r2.view(np.uint64))
    rows = np.diff(one[0])
    assert rows.max() > 0


def test_iteration_stays_in_the_range_of_p(case):
    ctx = case["ctx"]
    t, valid = forward_and_system(case, crit_chi=1.0)
    P, nodes, n_vel_m = mesh_p(ctx)
    v0 = ctx.model(0).copy()
    z = {k: ctx.model(k).copy() for k in IFACES}
    dm, d = ctx.damped_update(lam_damp=0.0, smooth_type=1, w_smooth=0.3, w_smooth_iface=0.3)
    vmask = np.ascontiguousarray(valid, dtype=np.int32)
    dm_step, r = ctx.line_search(case["t_obs"], dm, crit_chi=1.0, valid=vmask)
    print(f"mesh iteration: rms {r.rms_orig:.5f} -> {r.rms_new:.5f}, step {r.step}", flush=True)
    assert r.step > 0.0 and r.rms_new < r.rms_orig
    v1 = ctx.model(0)
    idx = np.asarray(ctx.param_index())
    ds = np.where(idx > 0, 1.0 / v1 - 1.0 / v0, 0.0)
    # the slowness change at the mesh nodes, prolonged, is the change everywhere
    ds_mesh = np.array([ds[iz, ix] for (ix, iz) in nodes])
    fine = [(ix, iz) for iz in range(idx.shape[0]) for ix in range(idx.shape[1]) if idx[iz, ix] > 0]
    pred = P[:len(fine), :n_vel_m] @ ds_mesh
    got = np.array([ds[iz, ix] for (ix, iz) in fine])
    # nodes where a clamp acts are left out: below 2 km/s, 30 % from the start
    ok = np.array([v0[iz, ix] >= 2.0 and 0.7 * v0[iz, ix] < v1[iz, ix] < 1.3 * v0[iz, ix] and 2.0 < v1[iz, ix] < 10.0
                   for (ix, iz) in fine])
    assert np.abs(got - pred)[ok].max() <= 1e-12 * max(np.abs(got).max(), 1e-30)
    assert np.abs(got).max() > 0.0
    # the accepted trial's fields are the committed model's: the next forward
    # of the same sources reuses them, and its rays equal those of a full solve
    ctx.forward(case["obs_x"])
    assert ctx.stats().launches == 0
    t_reused, v_reused = ctx.trace(*case["rays"])
    ctx.model_put(0, ctx.model(0))               # any model change: the fields are solved again
    ctx.forward(case["obs_x"])
    assert ctx.stats().launches > 0
    t_full, v_full = ctx.trace(*case["rays"])
    assert np.array_equal(t_reused.view(np.uint64), t_full.view(np.uint64))
    assert np.array_equal(v_reused, v_full)
    for k in IFACES:
        ctx.model_put(k, z[k])
    ctx.model_put(0, v0)


def test_invert2d_program_on_the_mesh(case, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("invert2d_hip not built (python -m fmt2d_amd.build)")
    from fmt2d_amd import formats, synth
    d0 = case["dir0"]
    geom = formats.read_geom(os.path.join(d0, "geom.dat"))
    t_obs = case["t_obs"]
    m_true, m0, g = synth.inversion_case(dx=1.0, phases=True)
    synth.write_fmt2d_case(str(tmp_path), m0, synth.with_picks(g, np.where(t_obs >= 0, t_obs, 0.0), t_obs >= 0))
    with open(tmp_path / "frechet.in", "w") as f:
        f.write("1\n1\n1\n1\n2\n3 4\n")
    with open(tmp_path / "invert2d.in", "w") as f:
        f.write("model.smesh\nmoho.dat\nsed.dat\ngeom.dat\n")
        for v in [4, 0.0, 0.3, 1.0, 400, 0.05, 0.1, 0.3, 1.0, 1.0, 1.0, 1.0, 10.0, 3.0]:
            f.write(f"{v}\n")
        f.write("2 0.2 0.5 0.3\n-1\n8.0\n0.0\n5.0 10.0\n")
    print("[gpu] invert2d_hip on the 5 km / 10 km mesh", flush=True)
    r = subprocess.run([EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "Inversion mesh strides (velocity, interfaces):    5  10" in r.stdout
    its = [ln for ln in r.stdout.splitlines() if ln.startswith("Iteration")]
    print("\n".join(its), flush=True)
    rows = [ln.replace("=", "= ").split() for ln in open(tmp_path / "invert_log_latest.txt") if ln.startswith("Iteration")]
    assert len(rows) == 4
    rms0 = float(rows[0][rows[0].index("RMS=") + 1])
    rms_last = float(rows[-1][rows[-1].index("RMS_new=") + 1])
    assert rms_last < 0.8 * rms0
