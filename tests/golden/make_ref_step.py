"""Writes tests/golden/ref_step.npz: one whole iteration of the reference
inversion (oracle/_ref/invert2d_dump, built from the reference sources by
oracle/build_ref.sh) with automatic damping and the line search, on two
cases of fmt2d_amd.synth.inversion_case:

  pg:    Pg only, velocity grid inverted (velocity damping search);
  iface: all five phases, velocity grid + sediment base + Moho inverted,
         joint update (interface and velocity damping searches, interface
         steps and the geometry rebuild).

Picks are the reference's receiver times on the true model plus seeded
noise (fmt2d_amd.synth.with_picks).  Kept per case: the picks, the damping
weights and mean updates the search chose, the accepted step, the RMS before
and after, the next ladder index, the step-scaled update and the model the
iteration left (velocity grid, inverted interfaces' depths).  Run from the
repository root:

    python tests/golden/make_ref_step.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fmt2d_amd import formats, synth  # noqa: E402
from oracle.raydump import read_raydump  # noqa: E402

# niter, lam_damp (0: automatic), lam_smooth, crit_chi, lsqr_max_iter,
# target_pert, lam_damp_iface, lam_smooth_iface, refl_weight, phase weights
# (PmP, Pg_ss, PsedP), corr_x, corr_z
INV_IN = [1, 0.0, 0.3, 1.0, 400, 0.05, 0.1, 0.3, 1.0, 1.0, 1.0, 1.0, 10.0, 3.0]
# alt: alternating mode of the iteration (0 joint); tdd: target mean
# interface update (0.05 km, below the undamped one, so its search runs)
CASES = {"pg": dict(phases=False, ifaces=(), alt=-1, tdd=0.0),
         "iface": dict(phases=True, ifaces=(3, 4), alt=0, tdd=0.05)}


def write_inputs(d, ifaces, alt, tdd):
    with open(os.path.join(d, "frechet.in"), "w") as f:
        f.write("1\n1\n1\n1\n")
        f.write(f"{len(ifaces)}\n" + (" ".join(str(k) for k in ifaces) + "\n" if ifaces else ""))
    with open(os.path.join(d, "invert2d.in"), "w") as f:
        f.write("model.smesh\nmoho.dat\nsed.dat\ngeom.dat\n")
        for v in INV_IN:
            f.write(f"{v}\n")
        f.write(f"1 {tdd} 0.5 0.3\n{alt}\n")


def read_step_dump(path):
    with open(path, "rb") as f:
        nparm, ls_idx = (int(x) for x in np.fromfile(f, dtype=np.int32, count=2))
        sc = np.fromfile(f, dtype=np.float64, count=7)
        dm = np.fromfile(f, dtype=np.float64, count=nparm)
        vnx, vnz = (int(x) for x in np.fromfile(f, dtype=np.int32, count=2))
        vel = np.fromfile(f, dtype=np.float64, count=vnx * vnz).reshape(vnz, vnx)
        nif = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        zs = []
        for _ in range(nif):
            n = int(np.fromfile(f, dtype=np.int32, count=1)[0])
            zs.append(np.fromfile(f, dtype=np.float64, count=n))
    return dict(ls_idx=ls_idx, scal=sc, dm=dm, vel=vel, ifz=zs)


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "invert2d_dump")
    if not os.path.exists(exe):
        raise SystemExit("build the reference first: bash oracle/build_ref.sh")
    out = {}
    for name, c in CASES.items():
        m_true, m0, g = synth.inversion_case(phases=c["phases"])
        with tempfile.TemporaryDirectory() as d:
            gp = synth.with_picks(g, np.ones(g.nrec), np.ones(g.nrec, bool))
            synth.write_fmt2d_case(d, m_true, gp)
            write_inputs(d, c["ifaces"], c["alt"], c["tdd"])
            env = dict(os.environ, FMT2D_RAYDUMP="raydump.txt")
            subprocess.run([exe], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL)
            rays = read_raydump(os.path.join(d, "raydump.txt"))
        t_true = np.array([r["t"] for r in rays])
        v_true = np.array([r["valid"] for r in rays])
        gi = synth.with_picks(g, t_true, v_true)
        with tempfile.TemporaryDirectory() as d:
            synth.write_fmt2d_case(d, m0, gi)
            write_inputs(d, c["ifaces"], c["alt"], c["tdd"])
            env = dict(os.environ, FMT2D_STEPDUMP="step.bin")
            r = subprocess.run([exe], cwd=d, env=env, check=True, capture_output=True, text=True)
            S = read_step_dump(os.path.join(d, "step.bin"))
            t_obs = formats.read_geom(os.path.join(d, "geom.dat")).t_obs
        out[name + "_t_obs"] = t_obs
        out[name + "_scal"] = S["scal"]          # wd, wd_iface, avg_dmv, avg_dmd, step, rms_orig, rms_new
        out[name + "_ls_idx"] = np.array([S["ls_idx"]])
        out[name + "_dm"] = S["dm"]
        out[name + "_vel"] = S["vel"]
        for k, z in enumerate(S["ifz"]):
            out[f"{name}_ifz{k}"] = z
        log = [ln for ln in r.stdout.splitlines() if "Final:" in ln or "step=" in ln]
        print(name, S["scal"], "ls_idx", S["ls_idx"], "|", " / ".join(s.strip() for s in log))
    path = os.path.join(ROOT, "tests", "golden", "ref_step.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
