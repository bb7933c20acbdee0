"""32-bit field storage of the tile solver (fmt2d_set_field_precision 32)
on the 0.1 km grid (2001x301), three OBS of the 60-OBS multistage case, all
twelve stage fields: grid-node times stored in 32-bit fixed point (2^-24 s),
raw gradients in FP32, every update computed in FP64.

Against the FP64 path (fmt2d_field_snapshot / fmt2d_field_compare):

* every field reaches exactly the same slots;
* almost every node moves by the rounding of its stored time (below
  3e-5 s); where the rounded times tip a node's second-order switch (the
  gradient test of regular_update2d) or a tie the other way, the node and
  the nodes downstream of it take a different update, a local cluster of a
  few hundred nodes with a few 1e-4 s.  Bars: nodes beyond north_star's
  1e-4 s at most 1e-4 of the reached slots, largest |dT| below 1e-3 s,
  every field with a node beyond 1e-4 s named (measured here: none, worst
  2.4e-5 s; on all 60 OBS, the bench's fp32_storage figure: 2 of 720
  fields, 1074 of 110.7 M reached slots, worst 3.0e-4 s);
* two 32-bit forwards give the same bits (every node is rounded to the
  same fixed-point grid, whatever the order of the updates);
* the adjoint-state gradient refuses an FP32 forward (its re-evaluation
  needs the FP64 fixed point)."""
import os

import numpy as np
import pytest

from conftest import load_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def f32(tmp_path_factory):
    from fmt2d_amd import hip, synth
    m, g = synth.multistage_phases_case()
    d = str(tmp_path_factory.mktemp("f32"))
    synth.write_fmt2d_case(d, m, g)
    ctx = hip.Context(0)
    ctx.set_model(*load_case(d))
    obs = list(g.obs_x)
    print(f"[gpu] FP32 field storage, 0.1 km grid, OBS {obs}", flush=True)
    ctx.forward(obs)
    ctx.field_snapshot()
    ctx.set_field_precision(32)
    ctx.forward(obs)
    yield ctx, obs
    ctx.set_field_precision(64)
    ctx.close()


def test_fp32_fields_against_fp64(f32):
    ctx, obs = f32
    mdt, nd, nover = ctx.field_compare(1e-4)
    summ = ctx.summary()
    per = mdt.reshape(len(obs), 12)
    print("max |dT| per stage field:", [f"{x:.1e}" for x in per.max(axis=0)])
    print("fields with nodes beyond 1e-4 s (OBS, tf, nodes, max s):",
          [(int(k // 12), int(k % 12 + 1), int(nover[k]), float(mdt[k])) for k in np.flatnonzero(nover > 0)])
    assert np.all(nd == 0)
    assert mdt.max() < 1e-3
    assert np.sum(nover) <= 1e-4 * np.sum(summ["reached"])


def test_fp32_forward_is_deterministic(f32):
    ctx, obs = f32
    ctx.field_snapshot()
    ctx.forward(obs)
    mdt, nd, _ = ctx.field_compare()
    assert np.all(mdt == 0.0) and np.all(nd == 0)


def test_adjoint_refuses_fp32_forward(f32):
    from fmt2d_amd import hip
    ctx, obs = f32
    ctx.trace([0], [1], [obs[0] + 5.0], [0.01])
    with pytest.raises(hip.Fmt2dError, match="FP64 field storage"):
        ctx.adjoint_gradient(np.ones(1))
