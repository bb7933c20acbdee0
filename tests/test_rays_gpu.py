"""HIP ray tracer and Frechet rows vs the reference's own rays.

The fixture (tests/golden/ref_rays_multiphase.npz, written by
tests/golden/make_ref_rays.py from the reference inversion built from its
sources) holds 848 rays of all five phases from three OBS of the multiphase
case: validity, receiver time, points per section, full point lists and
Frechet rows (velocity grid + sediment base + Moho).  Bar: the ray sparsity
pattern identical, Frechet entries within 1e-3 relative and receiver times
within 1e-4 s; the path reproduces them bit for bit and the test says so.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "ref_rays_multiphase.npz")


@pytest.fixture(scope="module")
def traced(tmp_path_factory):
    from fmt2d_amd import formats, hip, synth
    Z = np.load(FIX)
    d = str(tmp_path_factory.mktemp("mp"))
    m, g = synth.multiphase_case()
    synth.write_fmt2d_case(d, m, g)
    sm = formats.read_smesh(os.path.join(d, "model.smesh"))
    sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
    mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
    geom = formats.read_geom(os.path.join(d, "geom.dat"))
    ctx = hip.Context(0)
    ctx.set_model(sm, sx, sz, mx, mz)
    ctx.set_inversion(True, [3, 4])
    srcs = sorted(set(Z["src"].tolist()))
    local = {s: k for k, s in enumerate(srcs)}
    print(f"[gpu] forward multiphase case, sources {[geom.obs_x[s] for s in srcs]}", flush=True)
    ctx.forward([geom.obs_x[s] for s in srcs])
    rcv = Z["rcv"]
    src = [local[s] for s in Z["src"]]
    t, v = ctx.trace(src, geom.ray_type[rcv], geom.rec_x[rcv], geom.rec_z[rcv])
    rp, col, val = ctx.frechet_csr(index_base=1)
    return ctx, Z, t, v, rp, col, val


def test_receiver_times_and_validity(traced):
    ctx, Z, t, v, *_ = traced
    assert np.array_equal(v, Z["valid"])
    ok = Z["valid"]
    dt = np.abs(t[ok] - Z["t"][ok])
    assert dt.max() < 1e-4
    print(f"rays {len(v)}, valid {ok.sum()}, receiver times bit-identical: {np.mean(t[ok] == Z['t'][ok]):.4f}")
    assert np.array_equal(t[ok].view(np.int64), Z["t"][ok].view(np.int64))


def test_points_per_section_and_paths(traced):
    ctx, Z, t, v, *_ = traced
    ok = np.where(Z["valid"])[0]
    for k in ok:
        assert np.array_equal(ctx.ray_info(int(k))["npts"], Z["npts"][k]), f"ray {k}: points per section differ"
    n = 0
    for key in Z.files:
        if not key.startswith("pts_"):
            continue
        _, k, s = key.split("_")
        x, z = ctx.ray_points(int(k), int(s))
        ref = Z[key]
        assert np.array_equal(np.stack([x, z]), ref), f"ray {k} section {s}: points differ"
        n += 1
    assert n > 50


def test_frechet_rows(traced):
    ctx, Z, t, v, rp, col, val = traced
    rrp, rcol, rval = Z["rowptr"], Z["col"], Z["val"]
    nbit = 0
    for k in range(len(v)):
        a, b = rp[k] - 1, rp[k + 1] - 1
        ra, rb = rrp[k], rrp[k + 1]
        if not Z["valid"][k]:
            assert b == a
            continue
        assert np.array_equal(col[a:b], rcol[ra:rb]), f"ray {k}: Frechet sparsity differs"
        rel = np.abs(val[a:b] - rval[ra:rb]) / np.maximum(np.abs(rval[ra:rb]), 1e-30)
        assert rel.max(initial=0.0) < 1e-3, f"ray {k}: Frechet entry off by {rel.max():.2e}"
        nbit += int(np.sum(val[a:b] == rval[ra:rb]))
    print(f"Frechet entries {len(rcol)}, bit-identical {nbit}")
    assert nbit == len(rcol)


FINE = os.path.join(ROOT, "tests", "golden", "ref_rays_fine.npz")


@pytest.fixture(scope="module")
def traced_fine(tmp_path_factory):
    """The 0.1 km grid (2001x301), three OBS of the 60-OBS multistage case,
    all five phases (tests/golden/ref_rays_fine.npz, written by
    tests/golden/make_ref_rays_fine.py from the reference inversion)."""
    from fmt2d_amd import formats, hip, synth
    Z = np.load(FINE)
    d = str(tmp_path_factory.mktemp("fine"))
    m, g = synth.multistage_phases_case()
    synth.write_fmt2d_case(d, m, g)
    sm = formats.read_smesh(os.path.join(d, "model.smesh"))
    sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
    mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
    geom = formats.read_geom(os.path.join(d, "geom.dat"))
    ctx = hip.Context(0)
    ctx.set_model(sm, sx, sz, mx, mz)
    ctx.set_inversion(True, [3, 4])
    print(f"[gpu] forward 0.1 km grid, sources {list(geom.obs_x)}, {len(geom.rec_x)} rays", flush=True)
    ctx.forward(list(geom.obs_x))
    t, v = ctx.trace(geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
    rp, col, val = ctx.frechet_csr(index_base=1)
    yield ctx, Z, t, v, rp, col, val
    ctx.close()


def test_fine_grid_validity_and_times(traced_fine):
    """Which rays exist at 0.1 km (the 5000-point section cap, head-wave and
    interface rules) and their receiver times, every ray, bit for bit."""
    ctx, Z, t, v, *_ = traced_fine
    assert np.array_equal(v, Z["valid"]), f"{int(np.sum(v != Z['valid']))} rays differ in validity"
    ok = Z["valid"]
    assert np.array_equal(t[ok].view(np.int64), Z["t"][ok].view(np.int64)), \
        f"max |dt| {np.max(np.abs(t[ok] - Z['t'][ok])):.3e}"
    print(f"rays {len(v)}, valid {int(ok.sum())}; invalid per phase:",
          {p: int(np.sum(~ok & (Z["path"] == p))) for p in range(1, 6)})
    assert 0 < ok.sum() < len(ok)


def test_fine_grid_points(traced_fine):
    """Points per section of every valid ray identical; the point lists
    within 1e-12 km (bit for bit but for the last bit where a step reads the
    refit gradient of a sea-surface node of a water-layer stage, whose
    least-squares fit rounds differently in its last bit: 1e-17 km)."""
    ctx, Z, t, v, *_ = traced_fine
    for k in np.flatnonzero(Z["valid"]):
        assert np.array_equal(ctx.ray_info(int(k))["npts"], Z["npts"][k]), f"ray {k}: points per section differ"
    n, npt, nbit, dmax = 0, 0, 0, 0.0
    for key in Z.files:
        if not key.startswith("pts_"):
            continue
        _, k, s = key.split("_")
        x, z = ctx.ray_points(int(k), int(s))
        got, ref = np.stack([x, z]), Z[key]
        assert got.shape == ref.shape, f"ray {k} section {s}: {got.shape} points, reference {ref.shape}"
        d = np.abs(got - ref).max()
        assert d < 1e-12, f"ray {k} section {s}: points off by {d:.3e} km"
        dmax = max(dmax, d)
        npt += ref.shape[1]
        nbit += int(np.sum(np.all(got == ref, axis=0)))
        n += 1
    print(f"{n} sections, {npt} points, bit-identical {nbit}, max |d| {dmax:.1e} km")
    assert n > 10 and nbit > 0.99 * npt


def test_fine_grid_frechet_rows(traced_fine):
    """Frechet rows at 0.1 km: the sparsity pattern identical, entries within
    1e-9 relative (bar: 1e-3); rays whose points differ in the last bit (see
    above) carry that into their weights."""
    ctx, Z, t, v, rp, col, val = traced_fine
    rrp, rcol, rval = Z["rowptr"], Z["col"], Z["val"]
    nrow, nent, nbit, rmax = 0, 0, 0, 0.0
    for j, k in enumerate(Z["row_rays"]):
        a, b = rp[k] - 1, rp[k + 1] - 1
        ra, rb = rrp[j], rrp[j + 1]
        if not Z["valid"][k]:
            assert b == a
            continue
        assert np.array_equal(col[a:b], rcol[ra:rb]), f"ray {k}: Frechet sparsity differs"
        rel = np.abs(val[a:b] - rval[ra:rb]) / np.maximum(np.abs(rval[ra:rb]), 1e-30)
        assert rel.max(initial=0.0) < 1e-9, f"ray {k}: Frechet entry off by {rel.max():.2e} relative"
        rmax = max(rmax, float(rel.max(initial=0.0)))
        nent += rb - ra
        nbit += int(np.sum(val[a:b] == rval[ra:rb]))
        nrow += 1
    print(f"{nrow} rows, {nent} entries, bit-identical {nbit}, max relative difference {rmax:.1e}")
    assert nrow > 10 and nbit > 0.95 * nent


def test_fine_grid_indexed_search_is_the_scan(traced_fine):
    """The interpolators' nearest-node fallback (interpolate_arrivaltime_reg,
    find_position_in_region2d) from the column index gives the reference
    scan's node: the same validity, times, entry points, point lists and
    Frechet rows for every ray of the 0.1 km fixture, bit for bit."""
    ctx, Z, t, v, rp, col, val = traced_fine
    n_idx = ctx.ray_nearest_calls()
    src = np.asarray(Z["src"]) if "src" in Z.files else None
    try:
        ctx.set_ray_search(1)
        geom_rays = ctx._last_trace_args
        t2, v2 = ctx.trace(*geom_rays)
        n_scan = ctx.ray_nearest_calls()
        rp2, col2, val2 = ctx.frechet_csr(index_base=1)
        assert np.array_equal(v, v2)
        assert np.array_equal(t.view(np.int64), t2.view(np.int64))
        assert np.array_equal(rp, rp2) and np.array_equal(col, col2)
        assert np.array_equal(val.view(np.int64), val2.view(np.int64))
    finally:
        ctx.set_ray_search(0)
    print(f"nearest-node searches per trace: {n_idx} (indexed), {n_scan} (scan)")
    assert n_idx == n_scan


def test_indexed_nearest_node_on_random_points(traced_fine):
    """fmt2d_ray_nearest_check: for 20 000 seeded points over and around the
    0.1 km model, in every region, the column index returns the scan's node
    (the first minimum in node order), and points on grid nodes and
    half-way between them (distance ties) are in the set."""
    ctx = traced_fine[0]
    info = ctx.info
    rng = np.random.default_rng(11)
    x0, x1 = info.x0 - 1.0, info.x0 + info.dx * (info.nx - 1) + 1.0
    z0, z1 = info.z0 - 1.0, info.z0 + info.dz * (info.nz - 1) + 1.0
    px = np.concatenate([rng.uniform(x0, x1, 16000), info.x0 + info.dx * rng.integers(0, info.nx, 4000) / 2.0])
    pz = np.concatenate([rng.uniform(z0, z1, 16000), info.z0 + info.dz * rng.integers(0, info.nz, 4000) / 2.0])
    checked = 0
    for region in range(1, 7):
        a, b = ctx.ray_nearest_check(region, px, pz)
        if np.all(a == -2):
            continue
        assert np.array_equal(a, b), f"region {region}: {int(np.sum(a != b))} points differ"
        checked += 1
    assert checked >= 5
