"""The Fortran host program fmt2d_amd/fortran/invert2d_hip against whole
runs of the reference's invert2d (tests/golden/ref_invert2d.npz, from the
reference built from its sources; see tests/golden/make_ref_invert2d.py):
the same input files in, every file the reference wrote compared byte for
byte (iteration log, initial and final residuals and ray density, the
updated .smesh, the inverted interfaces, the model_iter file).

Cases: Pg with the velocity grid inverted, three iterations of automatic
damping and Laplacian smoothing; all five phases with the velocity grid,
sediment base and Moho inverted under the default alternating schedule
(three velocity iterations, then an interface iteration whose line search
fails and reverts), Gaussian smoothing with interface smoothing, and the
best model restored at the end; all five phases with the velocity grid and
the seafloor inverted jointly (sources and the in-water test of receivers
stay on the seafloor of the set-up)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "ref_invert2d.npz")
EXE = os.path.join(ROOT, "fmt2d_amd", "fortran", "invert2d_hip")


def first_difference(a: bytes, b: bytes) -> str:
    la, lb = a.decode().splitlines(), b.decode().splitlines()
    for k, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return f"line {k + 1}:\n  ours: {x!r}\n  ref:  {y!r}"
    return f"line counts {len(la)} vs {len(lb)}"


@pytest.mark.parametrize("case", ["pg", "joint", "seafloor"])
def test_invert2d_files_match_reference(tmp_path, case):
    if not os.path.exists(EXE):
        pytest.fail("invert2d_hip not built (python -m fmt2d_amd.build)")
    Z = np.load(FIX)
    inputs = {k.split("/", 2)[2]: Z[k].tobytes() for k in Z.files if k.startswith(case + "/in/")}
    ref = {k.split("/", 2)[2]: Z[k].tobytes() for k in Z.files if k.startswith(case + "/out/")}
    for fn, data in inputs.items():
        (tmp_path / fn).write_bytes(data)
    print(f"[gpu] invert2d_hip, case {case}", flush=True)
    r = subprocess.run([EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    print("\n".join(ln for ln in r.stdout.splitlines() if ln.startswith("Iteration")), flush=True)
    wrote = sorted(set(os.listdir(tmp_path)) - set(inputs))
    assert wrote == sorted(ref), f"files written {wrote}, reference {sorted(ref)}"
    for fn, data in ref.items():
        ours = (tmp_path / fn).read_bytes()
        assert ours == data, f"{fn} differs from the reference's at {first_difference(ours, data)}"
