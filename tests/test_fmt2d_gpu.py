"""The Fortran forward program fmt2d_amd/fortran/fmt2d_hip against whole
runs of the reference's fmt2d (tests/golden/ref_fmt2d.npz, from the
reference built from its sources; see tests/golden/make_ref_fmt2d.py): the
same input files in, every file the reference wrote compared byte for byte
(arrivals, observed times, entry points, per-OBS ray files).

Cases on the checkerboard model at 1 km spacing, three OBS: receivers at
10 m depth (in the water: entry search, direct water wave for Pg) and 0.3 km
below the seafloor; phases assigned by the program from the offset, given
explicitly (1..5), and a sparse line with the ray files written; explicit
phases on a model whose sediment base and Moho carry random Fourier
modes."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "ref_fmt2d.npz")
EXE = os.path.join(ROOT, "fmt2d_amd", "fortran", "fmt2d_hip")


def first_difference(a: bytes, b: bytes) -> str:
    la, lb = a.decode().splitlines(), b.decode().splitlines()
    for k, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return f"line {k + 1}:\n  ours: {x!r}\n  ref:  {y!r}"
    return f"line counts {len(la)} vs {len(lb)}"


@pytest.mark.parametrize("case", ["auto", "phases", "rays", "fourier"])
def test_fmt2d_files_match_reference(tmp_path, case):
    if not os.path.exists(EXE):
        pytest.fail("fmt2d_hip not built (python -m fmt2d_amd.build)")
    Z = np.load(FIX)
    inputs = {k.split("/", 2)[2]: Z[k].tobytes() for k in Z.files if k.startswith(case + "/in/")}
    ref = {k.split("/", 2)[2]: Z[k].tobytes() for k in Z.files if k.startswith(case + "/out/")}
    for fn, data in inputs.items():
        (tmp_path / fn).write_bytes(data)
    print(f"[gpu] fmt2d_hip, case {case}", flush=True)
    r = subprocess.run([EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    wrote = sorted(set(os.listdir(tmp_path)) - set(inputs))
    assert wrote == sorted(ref), f"files written {wrote}, reference {sorted(ref)}"
    bad = []
    for fn, data in sorted(ref.items()):
        ours = (tmp_path / fn).read_bytes()
        if ours != data:
            bad.append(f"{fn} differs from the reference's at {first_difference(ours, data)}")
    assert not bad, "\n".join(bad)
