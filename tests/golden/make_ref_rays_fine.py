"""Writes tests/golden/ref_rays_fine.npz: rays and Frechet rows of the
reference inversion (oracle/_ref/invert2d_dump, built from the reference
sources by oracle/build_ref.sh) on the 0.1 km grid (2001x301) of the full
multistage forward, OBS 0, 30 and 59 of its 60 (x = 2.0, 101.0, 196.7 km),
1000 shots every 0.2 km, receiver lines of all five phases in the
reference's windows (fmt2d_amd.synth.multistage_phases_case); velocity grid
plus sediment base and Moho inverted.  At 0.1 km the 5000-point cap of a ray
section (rays2d.f90:206, 320-324) decides which long rays exist.

Kept: for every ray its validity, receiver time and points per section; the
Frechet row of every 60th ray; the full point lists of the receiver lines
whose index within their OBS is a multiple of 150.  Run
from the repository root (a few minutes):

    python tests/golden/make_ref_rays_fine.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fmt2d_amd import synth  # noqa: E402
from oracle.raydump import read_raydump  # noqa: E402
from tests.golden.make_ref_rays import write_inversion_inputs  # noqa: E402

ROW_EVERY, PTS_EVERY = 60, 150


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "invert2d_dump")
    if not os.path.exists(exe):
        raise SystemExit("build the reference first: bash oracle/build_ref.sh")
    m, g = synth.multistage_phases_case()
    with tempfile.TemporaryDirectory() as d:
        synth.write_fmt2d_case(d, m, g)
        write_inversion_inputs(d)
        env = dict(os.environ, FMT2D_RAYDUMP="raydump.txt", FMT2D_RAYDUMP_EVERY=str(PTS_EVERY))
        subprocess.run([exe], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL)
        rays = read_raydump(os.path.join(d, "raydump.txt"))
    n = len(rays)
    assert n == g.nrec, (n, g.nrec)
    out = {"src": np.array([r["src"] - 1 for r in rays], dtype=np.int32),
           "path": np.array([r["path"] for r in rays], dtype=np.int32),
           "valid": np.array([r["valid"] for r in rays]),
           "t": np.array([r["t"] for r in rays])}
    npts = np.zeros((n, 4), dtype=np.int32)
    for k, r in enumerate(rays):
        for s in r["sections"]:
            npts[k, s["n"] - 1] = s["npts"]
    out["npts"] = npts
    rows = np.arange(0, n, ROW_EVERY)
    rp = [0]
    cols, vals = [], []
    for i in rows:
        cols.append(rays[i]["gi"]); vals.append(rays[i]["gv"]); rp.append(rp[-1] + len(rays[i]["gi"]))
    out["row_rays"] = rows.astype(np.int32)
    out["rowptr"] = np.array(rp, dtype=np.int64)
    out["col"] = np.concatenate(cols).astype(np.int32)
    out["val"] = np.concatenate(vals)
    # the dump lists the points of receiver lines whose index within their
    # OBS is a multiple of PTS_EVERY
    for k in range(n):
        if rays[k]["valid"]:
            for s in rays[k]["sections"]:
                if s["pts"]:
                    out[f"pts_{k}_{s['n']}"] = np.array(s["pts"]).T
    path = os.path.join(ROOT, "tests", "golden", "ref_rays_fine.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {n} rays ({int(out['valid'].sum())} valid), {len(rows)} Frechet rows with "
          f"{len(out['col'])} entries")


if __name__ == "__main__":
    main()
