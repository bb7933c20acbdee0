"""Writes tests/golden/ref_fmt2d.npz: whole runs of the reference's forward
program fmt2d (oracle/_ref/fmt2d, built from the reference sources by
oracle/build_ref.sh) on three cases, the checkerboard model of
fmt2d_amd.synth at 1 km grid spacing (201 x 31 nodes) and OBS at 45, 95
and 145 km:

  auto:   shots every 2 km at 10 m depth plus three receivers 0.3 km below
          the seafloor, ray type 0, so the program assigns Pg, PsedP, PmP
          and Pg_ss by offset;
  phases: shots every 1 km with explicit phases 1..5 in the reference's
          windows plus the buried receivers with Pg, PsedP and PmP;
  rays:   as auto with shots every 20 km, ray files written
          (write_rays = T);
  fourier: explicit phases 1..5 every 2 km on the model whose sediment base
          and Moho carry seeded random Fourier modes (fourier_seed 3).

Kept per case, as bytes: the input files and every file the run wrote
(arrivals, observed times, entry points, ray files).  Run from the
repository root:

    python tests/golden/make_ref_fmt2d.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fmt2d_amd import synth  # noqa: E402

DX = 1.0
OBS = (45.0, 95.0, 145.0)
BURIED_X = (20.5, 60.5, 130.5)
INPUTS = ["fmt2d.in", "model.smesh", "moho.dat", "sed.dat", "geom.dat"]


def buried(m, ray_types):
    z = np.interp(BURIED_X, m.xpos, m.topo) + 0.3
    return [(float(x), float(zz), rt, 0.0, 0.05) for x, zz in zip(BURIED_X, z) for rt in ray_types]


def case_geometry(name, m):
    if name == "fourier":
        g = synth.phase_windows_geometry(m, OBS, np.arange(1.0, 200.0, 2.0))
        for lst in g.receivers:
            lst.extend(buried(m, (1, 2, 3)))
        return g
    if name in ("auto", "rays"):
        shots = np.arange(1.0, 200.0, 2.0) if name == "auto" else np.arange(2.5, 200.0, 20.0)
        g = synth.make_geometry(OBS, shots, m, ray_types=(0,))
        extra = buried(m, (0,))
    else:
        shots = np.arange(0.5, 200.0, 1.0)
        g = synth.phase_windows_geometry(m, OBS, shots)
        extra = buried(m, (1, 2, 3))
    for lst in g.receivers:
        lst.extend(extra)
    return g


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "fmt2d")
    if not os.path.exists(exe):
        raise SystemExit("build the reference first: bash oracle/build_ref.sh")
    models = {"fourier": synth.make_model(dx=DX, fourier_seed=3)}
    out = {}
    for name, rays in (("auto", False), ("phases", False), ("rays", True), ("fourier", False)):
        m = models.get(name) or synth.make_model(dx=DX)
        g = case_geometry(name, m)
        with tempfile.TemporaryDirectory() as d:
            synth.write_fmt2d_case(d, m, g, write_rays=rays)
            before = set(os.listdir(d))
            subprocess.run([exe], cwd=d, check=True, stdout=subprocess.DEVNULL)
            wrote = sorted(set(os.listdir(d)) - before)
            sizes = {fn: os.path.getsize(os.path.join(d, fn)) for fn in wrote}
            print(f"{name}: wrote {sizes}")
            for fn in INPUTS:
                out[f"{name}/in/{fn}"] = np.frombuffer(open(os.path.join(d, fn), "rb").read(), np.uint8)
            for fn in wrote:
                out[f"{name}/out/{fn}"] = np.frombuffer(open(os.path.join(d, fn), "rb").read(), np.uint8)
    path = os.path.join(ROOT, "tests", "golden", "ref_fmt2d.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
