"""Writes tests/golden/ref_fields_dx2.npz: stage fields computed by the
reference's own modules (oracle/_ref/ref_fields, built from /root/reference
by oracle/build_ref.sh) on a coarse member of the synthetic model family
(2 km grid, 101x16 nodes), two sources, all twelve stages of run_forward2d
before the PmP head-wave post-pass.  Run from the repo root in a container
that has the reference mounted:  python tests/golden/make_ref_fields.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fmt2d_amd import synth  # noqa: E402
from oracle.refdump import read_refdump  # noqa: E402

SOURCES = [15.0, 133.0]


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_fields")
    if not os.path.exists(exe):
        raise SystemExit("oracle/_ref/ref_fields missing: run oracle/build_ref.sh")
    m = synth.make_model(dx=2.0)
    with tempfile.TemporaryDirectory() as d:
        synth.write_smesh(os.path.join(d, "model.smesh"), m)
        synth.write_interface(os.path.join(d, "moho.dat"), m.xpos, m.moho_z)
        synth.write_interface(os.path.join(d, "sed.dat"), m.xpos, m.sed_z)
        with open(os.path.join(d, "ref_fields.in"), "w") as f:
            f.write("model.smesh\nmoho.dat\nsed.dat\nout.bin\n")
            f.write(f"{len(SOURCES)}\n" + "".join(f"{s}\n" for s in SOURCES) + "12\n")
        subprocess.run([exe], cwd=d, check=True, capture_output=True)
        ref = read_refdump(os.path.join(d, "out.bin"), len(SOURCES), 12)
        files = {n: open(os.path.join(d, n)).read() for n in ("model.smesh", "moho.dat", "sed.dat")}
    out = {"sources": np.array(SOURCES), "nx": ref["nx"], "nz": ref["nz"],
           "fully_regular": ref["fully_regular"], "node_region": ref["node_region"]}
    for k, v in files.items():
        out["file_" + k.replace(".", "_")] = np.array(v)
    for r in range(1, 7):
        R = ref["regions"][r]
        for key in ("i1", "i2", "i3", "x", "z", "vel"):
            out[f"r{r}_{key}"] = R[key]
    for k in range(1, 6):
        for key, v in ref["isec"][k].items():
            out[f"i{k}_{key}"] = v
    for i, st in enumerate(ref["sources"]):
        seen = set()
        for itf, T, gx, gz in st:
            if itf in seen:
                continue
            seen.add(itf)
            out[f"s{i}_tf{itf}_T"] = T
            out[f"s{i}_tf{itf}_gx"] = gx
            out[f"s{i}_tf{itf}_gz"] = gz
    path = os.path.join(ROOT, "tests", "golden", "ref_fields_dx2.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
