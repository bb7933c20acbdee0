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
    python tests/golden/make_ref_lsqr.py --gauss   # ref_lsqr_gauss.npz (below)
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


# Gaussian smoothing cases: velocity grid + sediment base + Moho inverted,
# all five phases, correlation lengths 2 km (x), 1 km (z), 5 km (interfaces)
GAUSS_IN = [1, 0.1, 0.3, 1.0, 400, 0.05, 0.1, 0.3, 1.0, 1.0, 1.0, 1.0, 2.0, 1.0]
GAUSS_IFACES = (3, 4)
GAUSS_CORR_X_IFACE = 5.0


def write_gauss_inputs(d, smooth_type):
    with open(os.path.join(d, "frechet.in"), "w") as f:
        f.write("1\n1\n1\n1\n")
        f.write(f"{len(GAUSS_IFACES)}\n" + " ".join(str(k) for k in GAUSS_IFACES) + "\n")
    with open(os.path.join(d, "invert2d.in"), "w") as f:
        f.write("model.smesh\nmoho.dat\nsed.dat\ngeom.dat\n")
        for v in GAUSS_IN:
            f.write(f"{v}\n")
        f.write(f"{smooth_type} 0.0 0.5 0.3\n-1\n{GAUSS_CORR_X_IFACE}\n")


def row_sums(mat, n):
    """Front-to-back sum of each row's values (1-based CSR as dumped)."""
    rp, _, val = mat
    out = np.zeros(n)
    for i in range(n):
        a, b = rp[i] - 1, rp[i + 1] - 1
        out[i] = np.cumsum(val[a:b])[-1] if b > a else 0.0
    return out


def gauss_fixture(exe):
    """Writes tests/golden/ref_lsqr_gauss.npz: the first run_lsqr solve of the
    reference with smooth_type 2 (Gaussian R_h, R_v + interface R_if) and 3
    (Laplacian + interface R_if) on the all-phase small case."""
    m_true, m0, g = synth.inversion_case(phases=True)
    with tempfile.TemporaryDirectory() as d:
        gp = synth.with_picks(g, np.ones(g.nrec), np.ones(g.nrec, bool))
        synth.write_fmt2d_case(d, m_true, gp)
        write_gauss_inputs(d, 2)
        env = dict(os.environ, FMT2D_RAYDUMP="raydump.txt")
        subprocess.run([exe], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL)
        rays = read_raydump(os.path.join(d, "raydump.txt"))
    t_true = np.array([r["t"] for r in rays])
    v_true = np.array([r["valid"] for r in rays])
    gi = synth.with_picks(g, t_true, v_true)
    out = {}
    for stype in (2, 3):
        with tempfile.TemporaryDirectory() as d:
            synth.write_fmt2d_case(d, m0, gi)
            write_gauss_inputs(d, stype)
            env = dict(os.environ, FMT2D_LSQRDUMP="lsqr.bin")
            subprocess.run([exe], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL)
            D = read_lsqr_dump(os.path.join(d, "lsqr.bin"))
            t_obs = formats.read_geom(os.path.join(d, "geom.dat")).t_obs
        p = f"t{stype}_"
        # the system is the same for both smoothing types: kept once
        for k in ("rowptr", "gindx", "gval", "residual", "gtg", "dm_sum", "mvscale"):
            if k in out:
                assert np.array_equal(out[k], D[k]), k
            out[k] = D[k]
        out[p + "dm"] = D["dm"]
        out[p + "scal"] = np.array([D["ndata"], D["nparm"], D["nnz"], D["n_vel"], D["smooth_type"], D["max_iter"],
                                    D["iters"]], dtype=np.int64)
        out[p + "par"] = np.array([D["wdv"], D["wdv_iface"], D["w_smooth"], D["w_smooth_iface"], D["eps_min"],
                                   D["rnorm"], D["bnorm"], D["test2"], D["corr_x"], D["corr_z"],
                                   GAUSS_CORR_X_IFACE])
        if "rh" in D:
            for name in ("rh", "rv"):
                out[p + name + "_rowptr"] = D[name][0]
                out[p + name + "_rowsum"] = row_sums(D[name], D["nparm"])
        if "rif" in D:
            out[p + "rif_rowptr"], out[p + "rif_col"], out[p + "rif_val"] = D["rif"]
        out["t_obs"] = t_obs
        out["vx"], out["vz"], out["inv_idx"] = D["vx"], D["vz"], D["inv_idx"]
        out["iface_start"] = np.array([f["start"] for f in D["ifaces"]], dtype=np.int64)
        out["iface_id"] = np.array([f["id"] for f in D["ifaces"]], dtype=np.int64)
        for k, f in enumerate(D["ifaces"]):
            out[f"iface_x{k}"] = f["x"]
        print(f"smooth_type {stype}: ndata {D['ndata']}, nparm {D['nparm']}, nnz {D['nnz']}, "
              f"iterations {D['iters']}")
    path = os.path.join(ROOT, "tests", "golden", "ref_lsqr_gauss.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}")


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "invert2d_dump")
    if not os.path.exists(exe):
        raise SystemExit("build the reference first: bash oracle/build_ref.sh")
    if "--gauss" in sys.argv:
        gauss_fixture(exe)
        return
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
