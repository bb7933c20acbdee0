"""Writes tests/golden/ref_invert2d.npz: whole runs of the reference's
invert2d (oracle/_ref/invert2d, built from the reference sources by
oracle/build_ref.sh) on three cases of fmt2d_amd.synth.inversion_case at
1 km grid spacing (201 x 31 nodes):

  pg:    Pg only, velocity grid inverted, automatic damping, Laplacian
         smoothing, 3 iterations;
  joint: all five phases, velocity grid + sediment base + Moho inverted,
         the default alternating schedule (iterations 1-3 velocity only,
         4 interfaces only), Gaussian smoothing with interface smoothing,
         4 iterations;
  seafloor: all five phases, velocity grid + seafloor inverted jointly,
         hybrid smoothing, 2 iterations (sources and the in-water test of
         receivers stay on the seafloor of the set-up while it moves).

Picks are the reference's receiver times on the true model plus seeded
noise (fmt2d_amd.synth.with_picks).  Kept per case, as bytes: the input
files (invert2d.in, frechet.in, model, interfaces, geometry) and every file
the run wrote (log, initial/final residuals and ray density, the updated
.smesh, interfaces and model_iter file).  Run from the repository root:

    python tests/golden/make_ref_invert2d.py
"""
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fmt2d_amd import synth  # noqa: E402
from oracle.raydump import read_raydump  # noqa: E402

DX = 1.0
# invert2d.in after the four file names: niter, lam_damp (0: automatic),
# lam_smooth, crit_chi, lsqr_max_iter, target_pert, lam_damp_iface,
# lam_smooth_iface, refl_weight, phase weights (PmP, Pg_ss, PsedP), corr_x,
# corr_z; then "smooth_type target_dd iface_step_ratio max_rel_pert", the
# alternating mode and corr_x_iface
CASES = {
    "pg": dict(phases=False, ifaces=(),
               numbers=[3, 0.0, 0.3, 1.0, 400, 0.05, 0.1, 0.3, 1.0, 1.0, 1.0, 1.0, 10.0, 3.0],
               tail="1 0.0 0.5 0.3\n-1\n"),
    "joint": dict(phases=True, ifaces=(3, 4),
                  numbers=[4, 0.0, 0.3, 1.0, 400, 0.05, 0.1, 0.3, 1.0, 1.0, 1.0, 1.0, 10.0, 3.0],
                  tail="2 0.2 0.5 0.3\n-1\n8.0\n"),
    "seafloor": dict(phases=True, ifaces=(2,),
                     numbers=[2, 0.0, 0.3, 1.0, 400, 0.05, 0.1, 0.3, 1.0, 1.0, 1.0, 1.0, 10.0, 3.0],
                     tail="3 0.05 0.5 0.3\n0\n8.0\n"),
}
INPUTS = ["invert2d.in", "frechet.in", "model.smesh", "moho.dat", "sed.dat", "geom.dat"]


def write_inputs(d, c):
    with open(os.path.join(d, "frechet.in"), "w") as f:
        f.write("1\n1\n1\n1\n")
        ifs = c["ifaces"]
        f.write(f"{len(ifs)}\n" + (" ".join(str(k) for k in ifs) + "\n" if ifs else ""))
    with open(os.path.join(d, "invert2d.in"), "w") as f:
        f.write("model.smesh\nmoho.dat\nsed.dat\ngeom.dat\n")
        for v in c["numbers"]:
            f.write(f"{v}\n")
        f.write(c["tail"])


def main():
    ref = os.path.join(ROOT, "oracle", "_ref")
    dump, exe = os.path.join(ref, "invert2d_dump"), os.path.join(ref, "invert2d")
    if not (os.path.exists(dump) and os.path.exists(exe)):
        raise SystemExit("build the reference first: bash oracle/build_ref.sh")
    out = {}
    for name, c in CASES.items():
        m_true, m0, g = synth.inversion_case(dx=DX, phases=c["phases"])
        with tempfile.TemporaryDirectory() as d:
            synth.write_fmt2d_case(d, m_true, synth.with_picks(g, np.ones(g.nrec), np.ones(g.nrec, bool)))
            write_inputs(d, dict(c, numbers=[1] + c["numbers"][1:]))
            env = dict(os.environ, FMT2D_RAYDUMP="raydump.txt")
            subprocess.run([dump], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL)
            rays = read_raydump(os.path.join(d, "raydump.txt"))
        t_true = np.array([r["t"] for r in rays])
        v_true = np.array([r["valid"] for r in rays])
        with tempfile.TemporaryDirectory() as d:
            synth.write_fmt2d_case(d, m0, synth.with_picks(g, t_true, v_true))
            write_inputs(d, c)
            before = set(os.listdir(d))
            t0 = time.time()
            subprocess.run([exe], cwd=d, check=True, stdout=subprocess.DEVNULL)
            wrote = sorted(set(os.listdir(d)) - before)
            print(f"{name}: reference run {time.time() - t0:.1f} s, wrote {wrote}")
            for fn in INPUTS:
                out[f"{name}/in/{fn}"] = np.frombuffer(open(os.path.join(d, fn), "rb").read(), np.uint8)
            for fn in wrote:
                out[f"{name}/out/{fn}"] = np.frombuffer(open(os.path.join(d, fn), "rb").read(), np.uint8)
            with open(os.path.join(d, "invert_log_latest.txt")) as f:
                print(f.read().rstrip())
    path = os.path.join(ROOT, "tests", "golden", "ref_invert2d.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
