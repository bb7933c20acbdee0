"""Writes tests/golden/ref_headwave.npz: the PmP up-going stage field tf5 as
the reference inversion saves it after its head-wave post-pass
(run_forward2d, src/inversion2d.f90:4294-4409), dumped by the checker build
oracle/_ref/invert2d_dump (oracle/build_ref.sh, oracle/raydump.inc) on the
multiphase case's model (401x61 grid, 0.5 km) for two OBS, with the tf1 field
of the same forward beside it (before/after the post-pass is told apart by
tf5 alone; tf1 checks that the source and model are the oracle's).

Run from the repository root in a container with the reference mounted:

    python tests/golden/make_ref_headwave.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fmt2d_amd import synth  # noqa: E402
from tests.golden.make_ref_rays import write_inversion_inputs  # noqa: E402

OBS = [45.0, 145.0]


def read_field(path):
    with open(path, "rb") as f:
        n = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        a = np.fromfile(f, dtype=np.float64, count=3 * n)
        g = np.fromfile(f, dtype=np.float64, count=2 * n).reshape(n, 2)
    return a[:n], a[n:2 * n], a[2 * n:], g[:, 0].copy(), g[:, 1].copy()


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "invert2d_dump")
    if not os.path.exists(exe):
        raise SystemExit("build the reference first: bash oracle/build_ref.sh")
    m = synth.make_model(dx=0.5)
    shots = np.round(np.arange(0.5, 200.0 - 0.25, 5.0), 6)
    g = synth.phase_windows_geometry(m, OBS, shots)
    out = {"obs": np.array(OBS)}
    with tempfile.TemporaryDirectory() as d:
        synth.write_fmt2d_case(d, m, g)
        write_inversion_inputs(d)
        for k in range(len(OBS)):
            for itf in (1, 5):
                env = dict(os.environ, FMT2D_FIELDDUMP="field.bin", FMT2D_FIELDDUMP_SRC=str(k + 1),
                           FMT2D_FIELDDUMP_TF=str(itf))
                subprocess.run([exe], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL)
                x, z, T, gx, gz = read_field(os.path.join(d, "field.bin"))
                for key, v in (("T", T), ("gx", gx), ("gz", gz)):
                    out[f"s{k}_tf{itf}_{key}"] = v
        for name in ("model.smesh", "moho.dat", "sed.dat"):
            out["file_" + name.replace(".", "_")] = np.array(open(os.path.join(d, name)).read())
    path = os.path.join(ROOT, "tests", "golden", "ref_headwave.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
