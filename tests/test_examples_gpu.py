"""The checkerboard example workflows (fmt2d_amd/examples/checkerboard.py)
run end to end on the GPU host programs: every scenario's forward writes an
arrival for (nearly) every receiver line, and three inversion iterations
lower the RMS at every iteration, move the velocities toward the
checkerboard (positive correlation of recovered and true relative
perturbations; measured 0.25-0.46) and, in the joint scenarios, bring the
inverted interface closer to the true one."""
import os

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scenario", ["velocity", "moho", "sediment", "pgss"])
def test_checkerboard_scenario(tmp_path, scenario):
    from fmt2d_amd.examples import checkerboard as cb
    print(f"[gpu] checkerboard example: {scenario}", flush=True)
    rep = cb.main([scenario, "--workdir", str(tmp_path), "--niter", "3"])
    rms = rep["rms_by_iteration"]
    assert len(rms) == 4
    assert all(b < a for a, b in zip(rms, rms[1:])), rms
    assert rep["velocity_recovery"]["correlation"] > 0.2
    for k in cb.SCENARIOS[scenario].ifaces:
        e = rep[f"interface_{k}_rms_km"]
        assert e["final"] < e["start"]
    n_arr = sum(1 for _ in open(os.path.join(tmp_path, "obs_times.dat"))) - 1
    n_rec = sum(1 for ln in open(os.path.join(tmp_path, "geom.dat")) if ln.startswith("R"))
    assert n_arr >= 0.99 * n_rec
