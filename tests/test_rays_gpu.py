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
