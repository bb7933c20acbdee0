"""HIP eikonal path vs the oracle (FP64 FMM restatement) on the plain case
(401x61 grid) and on the multistage case (2001x301 grid).

Bar: the same reached-node set, traveltime within 1e-4 s max-abs on every
reached node of all twelve stage fields, gradient within 1e-3 relative at
99.9% of nodes.  (The path is built to reproduce the FP64 reference bit for
bit; the test prints the fraction of bitwise-equal nodes.)
"""
import numpy as np
import pytest

from conftest import load_case

pytestmark = pytest.mark.gpu

SOURCES = [5.0, 95.0, 195.0]


@pytest.fixture(scope="module")
def fields(plain_case):
    from fmt2d_amd import hip
    from oracle.pyoracle import OracleModel
    sm, sx, sz, mx, mz = load_case(plain_case)
    ctx = hip.Context(0)
    ctx.set_model(sm, sx, sz, mx, mz)
    print(f"[gpu] forward plain case sources={SOURCES}", flush=True)
    ctx.forward(SOURCES)
    om = OracleModel(sm, sx, sz, mx, mz)
    return ctx, om


@pytest.mark.parametrize("isrc", range(len(SOURCES)))
def test_all_stage_fields(fields, isrc):
    ctx, om = fields
    ref = om.forward_source(SOURCES[isrc], 12, headwave=True)
    worst, exact = [], []
    for itf in range(1, 13):
        T, gx, gz = ctx.field(isrc, itf)
        To, gxo, gzo = ref[itf - 1]
        fin = To < 1e19
        assert np.array_equal(fin, T < 1e19), f"tf{itf}: reached-node sets differ"
        dT = np.abs(T[fin] - To[fin])
        worst.append((itf, float(dT.max())))
        exact.append((itf, float(np.mean(T[fin] == To[fin]))))
        assert dT.max() < 1e-4, f"tf{itf}: max |dT| = {dT.max():.3e}"
        ok = fin & np.isfinite(gxo) & np.isfinite(gzo)
        g = np.hypot(gx[ok] - gxo[ok], gz[ok] - gzo[ok]) / np.maximum(np.hypot(gxo[ok], gzo[ok]), 1e-6)
        assert np.mean(g > 1e-3) < 1e-3, f"tf{itf}: {np.sum(g > 1e-3)} gradients off by >1e-3"
    print("max |dT| per stage:", worst)
    print("bitwise-equal fraction per stage:", exact)


MS_SOURCES = [2.0, 25.1, 77.9, 101.0, 196.7]


@pytest.fixture(scope="module")
def ms_fields(tmp_path_factory):
    from fmt2d_amd import hip, synth
    from oracle.pyoracle import OracleModel
    d = str(tmp_path_factory.mktemp("ms"))
    m, g = synth.multistage_case(dx=0.1, nobs=60)
    synth.write_fmt2d_case(d, m, g)
    sm, sx, sz, mx, mz = load_case(d)
    ctx = hip.Context(0)
    ctx.set_model(sm, sx, sz, mx, mz)
    print(f"[gpu] forward multistage case 2001x301 sources={MS_SOURCES}", flush=True)
    ctx.forward(MS_SOURCES)
    return ctx, OracleModel(sm, sx, sz, mx, mz)


@pytest.mark.parametrize("isrc", range(len(MS_SOURCES)))
def test_multistage_fields(ms_fields, isrc):
    """All twelve stage fields at 0.1 km, compared node by node: the reached
    sets and every traveltime bit for bit (the tolerance bar is 1e-4 s),
    gradients within 1e-6 relative (the refit gradients of the sea-surface
    nodes of the water stages, ~1% of their nodes, and of degenerate
    least-squares nodes differ in the last bits; the test prints how many)."""
    ctx, om = ms_fields
    ref = om.forward_source(MS_SOURCES[isrc], 12, headwave=True)
    for itf in range(1, 13):
        T, gx, gz = ctx.field(isrc, itf)
        To, gxo, gzo = ref[itf - 1]
        fin = To < 1e19
        assert np.array_equal(fin, T < 1e19), f"tf{itf}: reached-node sets differ"
        dT = np.abs(T[fin] - To[fin])
        assert dT.max() < 1e-4, f"tf{itf}: max |dT| = {dT.max():.3e}"
        assert np.array_equal(T[fin].view(np.int64), To[fin].view(np.int64)), \
            f"tf{itf}: not bit-identical (max |dT| {dT.max():.3e})"
        g = np.hypot(gx[fin] - gxo[fin], gz[fin] - gzo[fin]) / np.maximum(np.hypot(gxo[fin], gzo[fin]), 1e-6)
        assert g.max() < 1e-6, f"tf{itf}: gradient off by {g.max():.2e} relative"
        nb = int(np.sum((gx[fin] != gxo[fin]) | (gz[fin] != gzo[fin])))
        if nb:
            print(f"tf{itf}: {nb} of {fin.sum()} gradients differ in the last bits (max rel {g.max():.1e})")


@pytest.mark.parametrize("tuning", [(2.0, 6, 4, 12, (8, 8)), (3.0, 12, 2, 3, (8, 8)), (3.0, 16, 2, 16, (16, 8)),
                                    (2.0, 24, 1, 24, (8, 16)), (4.0, 8, 4, 8, (16, 8))])
def test_schedule_does_not_change_the_fields(fields, tuning):
    """The tile schedule (tile shape, window, sweep caps, in-tile sweep
    order) decides only how the fixed point is reached: every field of the
    plain case is the same, bit for bit, as with the default schedule."""
    ctx, _ = fields
    window, light, colors, heavy, tile = tuning
    before = [np.concatenate(ctx.field(i, f)).view(np.uint64).copy() for i in range(len(SOURCES)) for f in range(1, 13)]
    try:
        ctx.set_tile(*tile)
        ctx.set_tuning(window, light, colors)
        ctx.set_heavy_sweeps(heavy)
        ctx.forward(SOURCES)
        after = [np.concatenate(ctx.field(i, f)).view(np.uint64) for i in range(len(SOURCES)) for f in range(1, 13)]
    finally:
        ctx.set_tile()
        ctx.set_tuning()
        ctx.set_heavy_sweeps()
        ctx.forward(SOURCES)
    assert all(np.array_equal(a, b) for a, b in zip(before, after))


def test_level_launch_matches_the_oracle(fields):
    """The level launch mode (one cooperative dispatch per stage level, the
    solver's launches as phases between grid-wide barriers) on the plain
    case: every field bit for bit the default mode's, which the tests above
    hold to the oracle, and traveltimes of the first source bit for bit the
    oracle's.  Its launch cap gives the same error as the default mode's."""
    from fmt2d_amd import hip
    ctx, om = fields
    before = [np.concatenate(ctx.field(i, f)).view(np.uint64).copy() for i in range(len(SOURCES)) for f in range(1, 13)]
    try:
        ctx.set_solver_launch("level")
        ctx.forward(SOURCES)
        st = ctx.stats()
        print(f"level launch: {st.launches} launches in {st.dispatches} dispatches, {st.ms_total:.1f} ms")
        assert st.dispatches == 4
        after = [np.concatenate(ctx.field(i, f)).view(np.uint64) for i in range(len(SOURCES)) for f in range(1, 13)]
        assert all(np.array_equal(a, b) for a, b in zip(before, after))
        ref = om.forward_source(SOURCES[0], 12, headwave=True)
        for itf in range(1, 13):
            T = ctx.field(0, itf)[0]
            fin = ref[itf - 1][0] < 1e19
            assert np.array_equal(T[fin].view(np.int64), ref[itf - 1][0][fin].view(np.int64)), f"tf{itf}"
        ctx.set_launch_cap(3)
        with pytest.raises(hip.Fmt2dError, match="did not converge"):
            ctx.forward(SOURCES)
        ctx.set_launch_cap(0)
        ctx.forward(SOURCES)
        again = [np.concatenate(ctx.field(i, f)).view(np.uint64) for i in range(len(SOURCES)) for f in range(1, 13)]
        assert all(np.array_equal(a, b) for a, b in zip(before, again))
    finally:
        ctx.set_launch_cap(0)
        ctx.set_solver_launch("graph")
    with pytest.raises(ValueError):
        ctx.set_solver_launch("persistent")


def test_autotune_keeps_the_fields(fields):
    """fmt2d_autotune times the candidate schedules on this model, keeps the
    fastest and leaves the context solved: the fields are the default
    schedule's, bit for bit."""
    ctx, _ = fields
    before = [np.concatenate(ctx.field(i, f)).view(np.uint64).copy() for i in range(len(SOURCES)) for f in range(1, 13)]
    try:
        r = ctx.autotune(SOURCES, level=1)
        print(f"autotune: tile {r.tile_x}x{r.tile_z}, caps {r.max_sweeps}/{r.max_sweeps_heavy}, colours {r.colors}, "
              f"window {r.window_tiles}, {r.trials} forwards, {r.ms_start:.1f} -> {r.ms_best:.1f} ms")
        assert (r.tile_x, r.tile_z) in ((8, 8), (16, 8), (8, 16))
        assert r.trials >= 11 and r.ms_best <= r.ms_start
        after = [np.concatenate(ctx.field(i, f)).view(np.uint64) for i in range(len(SOURCES)) for f in range(1, 13)]
        assert all(np.array_equal(a, b) for a, b in zip(before, after))
    finally:
        ctx.set_tile()
        ctx.set_tuning()
        ctx.set_heavy_sweeps()
        ctx.forward(SOURCES)


def test_forward_after_launch_cap_matches_fresh(fields):
    """A forward that gives up at the launch cap (FMT2D_ENOCONV) leaves tiles
    marked active on the device; the next forward must start from clean tile
    state and give the same fields, bit for bit, as before."""
    from fmt2d_amd import hip
    ctx, _ = fields
    before = [np.concatenate(ctx.field(i, f)).view(np.uint64).copy() for i in range(len(SOURCES)) for f in range(1, 13)]
    try:
        ctx.set_launch_cap(3)
        with pytest.raises(hip.Fmt2dError, match="did not converge"):
            ctx.forward(SOURCES)
    finally:
        ctx.set_launch_cap(0)
    ctx.forward(SOURCES)
    after = [np.concatenate(ctx.field(i, f)).view(np.uint64) for i in range(len(SOURCES)) for f in range(1, 13)]
    assert all(np.array_equal(a, b) for a, b in zip(before, after))
