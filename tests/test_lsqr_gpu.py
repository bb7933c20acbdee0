"""HIP tomography system and LSQR vs the reference inversion.

The whole iteration runs on the GPU on the small inversion case
(fmt2d_amd.synth.inversion_case): forward + rays on the true model give the
receiver times the picks are made from, then forward + rays + Frechet rows on
the starting model, data selection/weighting and one LSQR solve.  The
fixture (tests/golden/ref_lsqr_small.npz) holds the reference's picks,
system and model update.  Bar: model update within 1e-4 relative; the path
reproduces gtg_diag, the residuals, the iteration count and the model update
bit for bit.  The parallel-reduction mode (fmt2d_set_lsqr_mode 1) stops up
to two iterations apart and its model update is held to 1e-2 relative (see
check_parallel_mode).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "ref_lsqr_small.npz")


def context_for(model, geom, d, ifaces=()):
    from fmt2d_amd import formats, hip, synth
    synth.write_fmt2d_case(d, model, geom)
    sm = formats.read_smesh(os.path.join(d, "model.smesh"))
    sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
    mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
    ctx = hip.Context(0)
    ctx.set_model(sm, sx, sz, mx, mz)
    ctx.set_inversion(True, list(ifaces))
    return ctx, formats.read_geom(os.path.join(d, "geom.dat"))


def test_inversion_step_matches_reference(tmp_path):
    from fmt2d_amd import synth
    Z = np.load(FIX)
    m_true, m0, g = synth.inversion_case()
    print("[gpu] true-model forward + rays (small inversion case)", flush=True)
    ctx, geom = context_for(m_true, g, str(tmp_path / "true"))
    ctx.forward(list(geom.obs_x))
    t, v = ctx.trace(geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
    assert np.array_equal(v, Z["valid_true"])
    assert np.array_equal(t[v], Z["t_true"][v])
    ctx.close()
    gi = synth.with_picks(g, t, v)
    print("[gpu] starting-model forward + rays + system + LSQR", flush=True)
    ctx, geom = context_for(m0, gi, str(tmp_path / "start"))
    assert np.array_equal(geom.t_obs, Z["t_obs"])
    ctx.forward(list(geom.obs_x))
    ctx.trace(geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
    nd, nnz, eps = ctx.build_system(geom.t_obs, crit_chi=1.0, refl_weight=1.0)
    ndata, nparm, nnz_r, n_vel, stype, maxit, iters = (int(x) for x in Z["scal"])
    assert (nd, nnz) == (ndata, nnz_r)
    assert eps == Z["par"][4]
    gtg, res = ctx.system()
    assert np.array_equal(gtg, Z["gtg"]) and np.array_equal(res, Z["residual"])
    dm, st = ctx.lsqr(wdv=0.1, wdv_iface=0.1, w_smooth=0.3, w_smooth_iface=0.3, max_iter=maxit)
    rel = np.max(np.abs(dm - Z["dm"])) / np.max(np.abs(Z["dm"]))
    print(f"LSQR iterations {st.iterations} (reference {iters}), model update max rel diff {rel:.1e}, "
          f"bit-identical {np.mean(dm == Z['dm']):.4f}, {st.ms:.1f} ms")
    assert rel < 1e-4
    assert st.iterations == iters
    assert np.array_equal(dm, Z["dm"])
    check_parallel_mode(ctx, Z["dm"], iters, dict(wdv=0.1, wdv_iface=0.1, w_smooth=0.3, w_smooth_iface=0.3,
                                                  max_iter=maxit))


def check_parallel_mode(ctx, ref, iters, kw):
    """fmt2d_set_lsqr_mode(1): double-double tree norms and wave-reduced
    products, rounded differently from the reference's ordered sums.  The
    Krylov recurrences amplify the difference: stopped at the same
    iteration the two modes' updates differ by 1e-3 to 5e-3 relative, and
    the stopping test fires up to two iterations apart (measured against the
    reference's update: 2.7e-4, 7.1e-3, 1.5e-3 on the three systems).  The
    bar here is 1e-2 relative and three iterations; north_star's 1e-4 holds
    only for the ordered mode, which is bit-exact."""
    try:
        ctx.set_lsqr_mode(1)
        dm1, st1 = ctx.lsqr(**kw)
    finally:
        ctx.set_lsqr_mode(0)
    rel = np.max(np.abs(dm1 - ref)) / np.max(np.abs(ref))
    n = min(st1.iterations, iters)
    kw_n = dict(kw, max_iter=n)
    dm0n, st0n = ctx.lsqr(**kw_n)
    try:
        ctx.set_lsqr_mode(1)
        dm1n, st1n = ctx.lsqr(**kw_n)
    finally:
        ctx.set_lsqr_mode(0)
    same = np.max(np.abs(dm1n - dm0n)) / np.max(np.abs(dm0n))
    print(f"parallel reductions: LSQR iterations {st1.iterations} (reference {iters}), model update max rel diff "
          f"{rel:.1e}, {st1.ms:.1f} ms; both modes stopped at {n} iterations: max rel diff {same:.1e}")
    assert rel < 1e-2
    assert abs(st1.iterations - iters) <= 3
    assert same < 1e-2


GAUSS = os.path.join(ROOT, "tests", "golden", "ref_lsqr_gauss.npz")


@pytest.mark.parametrize("stype", [2, 3])
def test_gaussian_smoothing_step_matches_reference(tmp_path, stype):
    """Gaussian (2) and hybrid (3) smoothing on the all-phase small case with
    the sediment base and the Moho inverted (tests/golden/ref_lsqr_gauss.npz):
    system and model update against the reference's, same bar."""
    from fmt2d_amd import synth
    Z = np.load(GAUSS)
    p = f"t{stype}_"
    ndata, nparm, nnz_r, n_vel, st_, maxit, iters = (int(x) for x in Z[p + "scal"])
    wdv, wdv_i, ws, ws_i, eps_r, rnorm, bnorm, test2, corr_x, corr_z, corr_xi = Z[p + "par"]
    m_true, m0, g = synth.inversion_case(phases=True)
    print(f"[gpu] all-phase case, smooth_type {stype}: forward + rays + system + LSQR", flush=True)
    ctx, geom = context_for(m0, g, str(tmp_path / "start"), ifaces=(3, 4))
    ctx.forward(list(geom.obs_x))
    ctx.trace(geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
    # the reference's first iteration with the alternating schedule (-1) is
    # velocity-only: interface columns scaled by 0
    nd, nnz, eps = ctx.build_system(Z["t_obs"], crit_chi=1.0, refl_weight=1.0, alt_mode=1)
    assert ctx.inversion_info() == (n_vel, nparm)
    assert (nd, nnz) == (ndata, nnz_r)
    assert eps == eps_r
    gtg, res = ctx.system()
    assert np.array_equal(gtg, Z["gtg"]) and np.array_equal(res, Z["residual"])
    dm, st = ctx.lsqr(wdv=wdv, wdv_iface=wdv_i, w_smooth=ws, w_smooth_iface=ws_i, smooth_type=stype,
                      max_iter=maxit, corr_x=corr_x, corr_z=corr_z, corr_x_iface=corr_xi)
    ref = Z[p + "dm"]
    rel = np.max(np.abs(dm - ref)) / np.max(np.abs(ref))
    print(f"smooth_type {stype}: LSQR iterations {st.iterations} (reference {iters}), model update max rel diff "
          f"{rel:.1e}, bit-identical {np.mean(dm == ref):.4f}, {st.ms:.1f} ms")
    assert rel < 1e-4
    assert st.iterations == iters
    assert np.array_equal(dm, ref)
    check_parallel_mode(ctx, ref, iters, dict(wdv=wdv, wdv_iface=wdv_i, w_smooth=ws, w_smooth_iface=ws_i,
                                              smooth_type=stype, max_iter=maxit, corr_x=corr_x, corr_z=corr_z,
                                              corr_x_iface=corr_xi))
