"""fmt2d_frechet_apply_noalloc (G x / G^T x staged in the library's fixed
device array) against fmt2d_frechet_apply (per-call device buffers): equal
bit for bit, both directions, on two traces in a row, the second with fewer
rays, so the second call runs over the array the first one left behind.
Also: the refusals (rows not built yet, wrong sizes)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx_case(tmp_path_factory):
    from fmt2d_amd import formats, hip, synth
    d = str(tmp_path_factory.mktemp("noalloc"))
    m, g = synth.multiphase_case(dx=1.0)
    synth.write_fmt2d_case(d, m, g)
    sm = formats.read_smesh(os.path.join(d, "model.smesh"))
    sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
    mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
    geom = formats.read_geom(os.path.join(d, "geom.dat"))
    print("[gpu] no-alloc G x: 1 km grid, all phases", flush=True)
    ctx = hip.Context(0)
    ctx.set_model(sm, sx, sz, mx, mz)
    ctx.set_inversion(True, [3, 4])
    ctx.forward(list(geom.obs_x))
    yield ctx, geom
    ctx.close()


def _check_trace(ctx, geom, keep, rng):
    from fmt2d_amd import hip
    ctx.trace((geom.obs_id - 1)[keep], geom.ray_type[keep], geom.rec_x[keep], geom.rec_z[keep])
    nray = int(np.sum(keep))
    n_par = ctx.mesh_info()[2]
    x = rng.normal(size=n_par)
    u = rng.normal(size=nray)
    # the rows of this trace are not built yet: the no-alloc call refuses
    with pytest.raises(hip.Fmt2dError, match="not built"):
        ctx.frechet_apply_noalloc(x, np.zeros(nray))
    ref_y = ctx.frechet_apply(x, transpose=False)
    y = ctx.frechet_apply_noalloc(x, np.zeros(nray))
    assert np.array_equal(y.view(np.uint64), ref_y.view(np.uint64))
    ref_g = ctx.frechet_apply(u, transpose=True)
    gt = ctx.frechet_apply_noalloc(u, np.zeros(n_par), transpose=True)
    assert np.array_equal(gt.view(np.uint64), ref_g.view(np.uint64))
    # again, over the array's leftovers
    y2 = ctx.frechet_apply_noalloc(x, np.full(nray, np.nan))
    assert np.array_equal(y2.view(np.uint64), ref_y.view(np.uint64))
    assert np.any(ref_y != 0.0) and np.any(ref_g != 0.0)
    return nray


def test_noalloc_matches_allocating_call_twice(ctx_case):
    ctx, geom = ctx_case
    rng = np.random.default_rng(7)
    n_all = len(geom.rec_x)
    first = _check_trace(ctx, geom, np.ones(n_all, dtype=bool), rng)
    keep = np.zeros(n_all, dtype=bool)
    keep[::3] = True
    second = _check_trace(ctx, geom, keep, rng)
    assert second < first


def test_noalloc_refuses_wrong_sizes(ctx_case):
    ctx, geom = ctx_case
    n_par = ctx.mesh_info()[2]
    with pytest.raises(ValueError):
        ctx.frechet_apply_noalloc(np.zeros(n_par + 1), np.zeros(ctx._nray))
