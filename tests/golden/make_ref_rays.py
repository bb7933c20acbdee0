"""Writes tests/golden/ref_rays_multiphase.npz: rays and Frechet rows of the
reference inversion (oracle/_ref/invert2d_dump, built from the reference
sources by oracle/build_ref.sh) on the multiphase case of
fmt2d_amd.synth.multiphase_case (401x61 grid, 20 OBS, all five phases;
velocity grid plus sediment base and Moho inverted).

Kept: every third receiver line of OBS 1, 8 and 15 (all phases): validity,
receiver time, points per section, the Frechet row, and the full point list
of every tenth kept ray.  Run from the repository root:

    python tests/golden/make_ref_rays.py
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

KEEP_OBS = (1, 8, 15)


def write_inversion_inputs(d, ifaces=(3, 4)):
    with open(os.path.join(d, "frechet.in"), "w") as f:
        f.write("1\n1\n1\n1\n")
        f.write(f"{len(ifaces)}\n" + (" ".join(str(k) for k in ifaces) + "\n" if ifaces else ""))
    with open(os.path.join(d, "invert2d.in"), "w") as f:
        f.write("model.smesh\nmoho.dat\nsed.dat\ngeom.dat\n")
        for v in [1, 0.1, 0.3, 1.0, 50, 0.05, 0.1, 0.3, 1.0, 1.0, 1.0, 1.0, 10.0, 3.0]:
            f.write(f"{v}\n")
        f.write("1 0.0 0.5 0.3\n-1\n")


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "invert2d_dump")
    if not os.path.exists(exe):
        raise SystemExit("build the reference first: bash oracle/build_ref.sh")
    m, g = synth.multiphase_case()
    with tempfile.TemporaryDirectory() as d:
        synth.write_fmt2d_case(d, m, g)
        write_inversion_inputs(d)
        env = dict(os.environ, FMT2D_RAYDUMP="raydump.txt", FMT2D_RAYDUMP_EVERY="1")
        subprocess.run([exe], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL)
        rays = read_raydump(os.path.join(d, "raydump.txt"))
    sel = [i for i, r in enumerate(rays) if r["src"] in KEEP_OBS and i % 3 == 0]
    out = {"rcv": np.array(sel, dtype=np.int32)}
    out["src"] = np.array([rays[i]["src"] - 1 for i in sel], dtype=np.int32)
    out["path"] = np.array([rays[i]["path"] for i in sel], dtype=np.int32)
    out["valid"] = np.array([rays[i]["valid"] for i in sel])
    out["t"] = np.array([rays[i]["t"] for i in sel])
    npts = np.zeros((len(sel), 4), dtype=np.int32)
    for k, i in enumerate(sel):
        for s in rays[i]["sections"]:
            npts[k, s["n"] - 1] = s["npts"]
    out["npts"] = npts
    rp = [0]
    cols, vals = [], []
    for i in sel:
        cols.append(rays[i]["gi"]); vals.append(rays[i]["gv"]); rp.append(rp[-1] + len(rays[i]["gi"]))
    out["rowptr"] = np.array(rp, dtype=np.int64)
    out["col"] = np.concatenate(cols).astype(np.int32)
    out["val"] = np.concatenate(vals)
    for k, i in enumerate(sel):
        if k % 10 == 0 and rays[i]["valid"]:
            for s in rays[i]["sections"]:
                if s["pts"]:
                    out[f"pts_{k}_{s['n']}"] = np.array(s["pts"]).T
    path = os.path.join(ROOT, "tests", "golden", "ref_rays_multiphase.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(sel)} rays, {len(out['col'])} Frechet entries")


if __name__ == "__main__":
    main()
