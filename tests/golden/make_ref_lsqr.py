"""Writes tests/golden/ref_lsqr_small.npz: the reference inversion's first
LSQR solve (oracle/_ref/invert2d_dump, built from the reference sources by
oracle/build_ref.sh) on the small inversion case of
fmt2d_amd.synth.inversion_case (401x61 grid, 3 OBS, shots every 1 km, Pg).

Picks are the reference's own receiver times on the true (checkerboard)
model plus seeded noise (fmt2d_amd.synth.with_picks, sigma 50 ms, seed 1);
the inversion starts from the model without the checkerboard with
lam_damp 0.1, lam_smooth 0.3, Laplacian smoothing, at most 400 iterations.
Kept: the picks, the system the reference hands to run_lsqr (CSR G after
weighting and scaling, residuals, gtg_diag, eps_min) and its result (model
update, iterations, rnorm, test2).  Run from the repository root:

    python tests/golden/make_ref_lsqr.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fmt2d_amd import formats, synth  # noqa: E402
from oracle.lsqr_oracle import read_lsqr_dump  # noqa: E402
from oracle.raydump import read_raydump  # noqa: E402

INV_IN = [1, 0.1, 0.3, 1.0, 400, 0.05, 0.1, 0.3, 1.0, 1.0, 1.0, 1.0, 10.0, 3.0]


def write_inputs(d):
    with open(os.path.join(d, "frechet.in"), "w") as f:
        f.write("1\n1\n1\n1\n0\n")
    with open(os.path.join(d, "invert2d.in"), "w") as f:
        f.write("model.smesh\nmoho.dat\nsed.dat\ngeom.dat\n")
        for v in INV_IN:
            f.write(f"{v}\n")
        f.write("1 0.0 0.5 0.3\n-1\n")


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "invert2d_dump")
    if not os.path.exists(exe):
        raise SystemExit("build the reference first: bash oracle/build_ref.sh")
    m_true, m0, g = synth.inversion_case()
    with tempfile.TemporaryDirectory() as d:
        # receiver times on the true model (first forward of the dump)
        gp = synth.with_picks(g, np.ones(g.nrec), np.ones(g.nrec, bool))
        synth.write_fmt2d_case(d, m_true, gp)
        write_inputs(d)
        env = dict(os.environ, FMT2D_RAYDUMP="raydump.txt")
        subprocess.run([exe], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL)
        rays = read_raydump(os.path.join(d, "raydump.txt"))
    t_true = np.array([r["t"] for r in rays])
    v_true = np.array([r["valid"] for r in rays])
    gi = synth.with_picks(g, t_true, v_true)
    with tempfile.TemporaryDirectory() as d:
        synth.write_fmt2d_case(d, m0, gi)
        write_inputs(d)
        env = dict(os.environ, FMT2D_LSQRDUMP="lsqr.bin")
        subprocess.run([exe], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL)
        D = read_lsqr_dump(os.path.join(d, "lsqr.bin"))
        t_obs = formats.read_geom(os.path.join(d, "geom.dat")).t_obs  # the picks as the reference read them
    out = {k: D[k] for k in ("rowptr", "gindx", "gval", "residual", "gtg", "dm_sum", "dm")}
    out.update(t_obs=t_obs, t_true=t_true, valid_true=v_true,
               scal=np.array([D["ndata"], D["nparm"], D["nnz"], D["n_vel"], D["smooth_type"], D["max_iter"],
                              D["iters"]], dtype=np.int64),
               par=np.array([D["wdv"], D["wdv_iface"], D["w_smooth"], D["w_smooth_iface"], D["eps_min"],
                             D["rnorm"], D["bnorm"], D["test2"]]))
    path = os.path.join(ROOT, "tests", "golden", "ref_lsqr_small.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: ndata {D['ndata']}, nparm {D['nparm']}, nnz {D['nnz']}, iterations {D['iters']}")


if __name__ == "__main__":
    main()
