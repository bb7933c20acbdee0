"""Outer-loop pieces of one inversion iteration vs the reference.

Automatic damping (secant searches over LSQR solves), clean-up and cap of the
update, the model step and the line search on the RMS after a forward, on
two cases (tests/golden/ref_step.npz, from the reference inversion built
from its sources; see tests/golden/make_ref_step.py): Pg with the velocity
grid inverted, and all five phases with the sediment base and the Moho
inverted jointly (both damping searches, interface steps, geometry
rebuild).  Bar: the update within 1e-4 relative; the path reproduces the
chosen damping weights, the mean updates, the accepted step, both RMS
values, the update and the updated model bit for bit.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "ref_step.npz")
CASES = {"pg": dict(phases=False, ifaces=(), alt=0, tdd=0.0), "iface": dict(phases=True, ifaces=(3, 4), alt=0,
                                                                              tdd=0.05)}


@pytest.mark.parametrize("name", ["pg", "iface"])
def test_iteration_outcome_matches_reference(tmp_path, name):
    from fmt2d_amd import formats, hip, synth
    Z = np.load(FIX)
    c = CASES[name]
    m_true, m0, g = synth.inversion_case(phases=c["phases"])
    d = str(tmp_path)
    synth.write_fmt2d_case(d, m0, g)
    sm = formats.read_smesh(os.path.join(d, "model.smesh"))
    sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
    mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
    geom = formats.read_geom(os.path.join(d, "geom.dat"))
    t_obs = Z[name + "_t_obs"]
    print(f"[gpu] {name}: forward + rays + system, damping search, line search", flush=True)
    ctx = hip.Context(0)
    ctx.set_model(sm, sx, sz, mx, mz)
    ctx.set_inversion(True, list(c["ifaces"]))
    ctx.forward(list(geom.obs_x))
    ctx.trace(geom.obs_id - 1, geom.ray_type, geom.rec_x, geom.rec_z)
    ctx.build_system(t_obs, crit_chi=1.0, refl_weight=1.0, alt_mode=c["alt"])
    dm, r = ctx.damped_update(lam_damp=0.0, lam_damp_iface=0.1, target_dv=0.05, target_dd=c["tdd"],
                              max_rel_pert=0.3, w_smooth=0.3, w_smooth_iface=0.3, smooth_type=1, max_iter=400)
    wd, wdi, admv, admd, step, rms0, rms1 = Z[name + "_scal"]
    print(f"damping: wd {r.wd} (ref {wd}), wd_iface {r.wd_iface} (ref {wdi}), {r.solves} LSQR solves; "
          f"avg_dmv {r.avg_dmv} (ref {admv}), avg_dmd {r.avg_dmd} (ref {admd})", flush=True)
    assert (r.wd, r.wd_iface) == (wd, wdi)
    assert r.avg_dmv == admv and r.avg_dmd == admd
    dm2, ls = ctx.line_search(t_obs, dm, crit_chi=1.0, iface_step_ratio=0.5, first_step=1)
    print(f"line search: step {ls.step} (ref {step}), RMS {ls.rms_orig} -> {ls.rms_new} (ref {rms0} -> {rms1}), "
          f"{ls.forwards} forwards", flush=True)
    assert ls.step == step and ls.rms_orig == rms0 and ls.rms_new == rms1
    assert ls.next_first_step == int(Z[name + "_ls_idx"][0])
    ref = Z[name + "_dm"]
    rel = np.max(np.abs(dm2 - ref)) / np.max(np.abs(ref))
    assert rel < 1e-4
    assert np.array_equal(dm2, ref)
    assert np.array_equal(ctx.model(0), Z[name + "_vel"])
    for k, itf in enumerate(c["ifaces"]):
        assert np.array_equal(ctx.model(itf), Z[f"{name}_ifz{k}"])
