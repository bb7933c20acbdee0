#!/usr/bin/env python3
"""Checkerboard resolution tests with the GPU host programs.

The workflows of the reference's example scripts (run_vel_checkerboard.py,
run_joint_checkerboard_moho.py, run_joint_checkerboard_sed.py,
run_pgss_checkerboard.py): build a starting model and a "true" model with a
+-5 % velocity checkerboard (and, for the joint cases, perturbed
interfaces), lay out OBS and receivers for the chosen phases, compute the
true model's times with ``fmt2d_hip``, add seeded Gaussian noise, invert
them with ``invert2d_hip`` and report how much of the checkerboard (and of
the interface perturbation) comes back.

Scenarios:

=========  =========================  =============================
name       phases                     inverted
=========  =========================  =============================
velocity   Pg (3-50 km), PmP (5-50)   velocity grid
moho       Pg, PmP (5-30 km)          velocity grid + Moho
sediment   Pg, PsedP (3-15 km)        velocity grid + sediment base
pgss       Pg, Pg_ss, Pg_ss2 (3-50)   velocity grid
=========  =========================  =============================

Usage::

    python -m fmt2d_amd.examples.checkerboard velocity --workdir /tmp/cb
    python -m fmt2d_amd.examples.checkerboard moho --niter 6 --dx 0.5

Every step writes the reference's text formats, so the same directory can be
handed to the reference's own fmt2d / invert2d.  The programs are built by
``python -m fmt2d_amd.build``.  Plots (PNG) are written when matplotlib is
importable; the text report is always printed.
"""
from __future__ import annotations

import argparse
import dataclasses
import os
import subprocess
import sys
from typing import Dict, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
sys.path.insert(0, os.path.dirname(PKG))

from fmt2d_amd import formats, synth  # noqa: E402

FMT2D = os.path.join(PKG, "fortran", "fmt2d_hip")
INVERT2D = os.path.join(PKG, "fortran", "invert2d_hip")

# phase ids of geom.dat (ray types)
PG, PSEDP, PMP, PGSS, PGSS2 = 1, 2, 3, 4, 5


@dataclasses.dataclass
class Scenario:
    phases: Dict[int, Tuple[float, float]]     # ray type -> (min, max) offset, km
    ifaces: Sequence[int]                        # interfaces inverted (3 sediment base, 4 Moho)
    niter: int
    lam_damp: float
    lam_smooth: float
    lam_damp_iface: float = 0.0
    lam_smooth_iface: float = 0.0
    max_rel_pert: float = 0.2


SCENARIOS = {
    "velocity": Scenario({PG: (3.0, 50.0), PMP: (5.0, 50.0)}, (), 8, 0.15, 0.3),
    "moho": Scenario({PG: (3.0, 50.0), PMP: (5.0, 30.0)}, (4,), 8, 0.25, 0.4, 0.1, 1.5, 0.15),
    "sediment": Scenario({PG: (3.0, 50.0), PSEDP: (3.0, 15.0)}, (3,), 8, 0.25, 0.4, 0.1, 1.5, 0.15),
    "pgss": Scenario({PG: (3.0, 50.0), PGSS: (3.0, 50.0), PGSS2: (3.0, 50.0)}, (), 8, 0.15, 0.3),
}

# model and survey of the examples (km): 120 km line, OBS every 8 km,
# receivers (shots) every 0.3 km at 5 m depth, checkerboard cells 18 x 5 km
XMAX, ZBOTTOM = 120.0, 25.0
OBS_X = np.arange(6.0, XMAX - 5.0, 8.0)
RCV_DX, RCV_Z = 0.3, 0.005
CHECKER = (0.05, 18.0, 5.0)
PICK_SIGMA = 0.03


def run(cmd, cwd, log):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    with open(os.path.join(cwd, log), "w") as f:
        f.write(r.stdout)
        f.write(r.stderr)
    if r.returncode != 0:
        raise SystemExit(f"{os.path.basename(cmd[0])} failed ({r.returncode}); see {os.path.join(cwd, log)}")
    return r.stdout


def interface_perturbation(x, amp, wavelength, phase=0.0):
    return amp * np.sin(2.0 * np.pi * x / wavelength + phase)


def models(name: str, dx: float):
    """(true, start) synth.Model pairs: the start has no checkerboard; in the
    joint scenarios the true interface also carries a sinusoid the start lacks."""
    amp, lx, lz = CHECKER
    true = synth.make_model(dx=dx, xmax=XMAX, z_bottom=ZBOTTOM, checker_amp=amp, checker_lx=lx, checker_lz=lz)
    start = synth.make_model(dx=dx, xmax=XMAX, z_bottom=ZBOTTOM, checker_amp=0.0)
    if name == "moho":
        true.moho_z = true.moho_z + interface_perturbation(true.xpos, 0.7, 40.0)
    if name == "sediment":
        true.sed_z = np.maximum(true.sed_z + interface_perturbation(true.xpos, 0.3, 25.0, 1.0), true.topo + 0.3)
    return true, start


def geometry(sc: Scenario, m) -> synth.Geometry:
    rx = np.round(np.arange(0.5, XMAX - 0.5 + 1e-9, RCV_DX), 6)
    recs = []
    for ox in OBS_X:
        lst = []
        for rt, (lo, hi) in sc.phases.items():
            off = np.abs(rx - ox)
            lst += [(float(x), RCV_Z, rt, 0.0, 0.05) for x in rx[(off >= lo) & (off <= hi)]]
        recs.append(lst)
    return synth.Geometry(obs_x=OBS_X, obs_z=np.interp(OBS_X, m.xpos, m.topo) + 0.05, receivers=recs)


def prepare(name: str, wd: str, dx: float) -> None:
    sc = SCENARIOS[name]
    true, start = models(name, dx)
    os.makedirs(wd, exist_ok=True)
    for tag, m in (("true", true), ("init", start)):
        synth.write_smesh(os.path.join(wd, f"{tag}_vp.smesh"), m)
        synth.write_interface(os.path.join(wd, f"{tag}_moho.dat"), m.xpos, m.moho_z)
        synth.write_interface(os.path.join(wd, f"{tag}_sed_bottom.dat"), m.xpos, m.sed_z)
    g = geometry(sc, true)
    synth.write_geom(os.path.join(wd, "geom.dat"), g)
    with open(os.path.join(wd, "fmt2d.in"), "w") as f:
        f.write("true_vp.smesh\ntrue_moho.dat\ntrue_sed_bottom.dat\ngeom.dat\nF\n"
                "arrivals.dat\nobs_times.dat\nrays_src\n")
    print(f"[1-3] models and geometry in {wd}: {len(OBS_X)} OBS, {g.nrec} receivers "
          f"({', '.join(f'type {k}' for k in sc.phases)})")


def forward(wd: str) -> None:
    run([FMT2D], wd, "fmt2d.log")
    n = sum(1 for _ in open(os.path.join(wd, "obs_times.dat"))) - 1
    print(f"[4] fmt2d_hip: {n} arrivals of the true model")


def picks(name: str, wd: str, seed: int = 1) -> int:
    """geom_inv.dat: the forward's times plus seeded noise for the scenario's
    phases (the reference scripts' pick file), -1 where there is none."""
    sc = SCENARIOS[name]
    obs = np.loadtxt(os.path.join(wd, "obs_times.dat"), skiprows=1, ndmin=2)
    lookup = {(int(r[0]), int(r[1]), int(r[2])): r[3] for r in obs}
    rng = np.random.default_rng(seed)
    geom = formats.read_geom(os.path.join(wd, "geom.dat"))
    n_ok = 0
    with open(os.path.join(wd, "geom_inv.dat"), "w") as f:
        f.write(f"{len(geom.obs_x)}\n")
        k = 0
        for s in range(len(geom.obs_x)):
            sel = np.flatnonzero(geom.obs_id == s + 1)
            f.write(f"S  {geom.obs_x[s]:.3f}  {geom.obs_z[s]:.5f}  {len(sel)}\n")
            for r in sel:
                k += 1
                rt = int(geom.ray_type[r])
                t = lookup.get((k, s + 1, rt))
                if rt in sc.phases and t is not None:
                    t = t + rng.normal(0.0, PICK_SIGMA)
                    n_ok += 1
                else:
                    t = -1.0
                f.write(f"R  {geom.rec_x[r]:.3f}  {geom.rec_z[r]:.3f}  {rt}  {t:.6f}  {PICK_SIGMA:.6f}\n")
    print(f"[5] {n_ok} picks with {1e3 * PICK_SIGMA:.0f} ms noise (seed {seed})")
    return n_ok


def invert(name: str, wd: str, niter: int = 0) -> str:
    sc = SCENARIOS[name]
    with open(os.path.join(wd, "frechet.in"), "w") as f:
        f.write("1\n1\n1\n1\n")                      # invert velocity grid 1 (P)
        f.write(f"{len(sc.ifaces)}\n")
        if sc.ifaces:
            f.write(" ".join(str(k) for k in sc.ifaces) + "\n")
    amp, lx, lz = CHECKER
    lines = ["init_vp.smesh", "init_moho.dat", "init_sed_bottom.dat", "geom_inv.dat", niter or sc.niter,
             sc.lam_damp, sc.lam_smooth, 0.5, 1000, 0.02, sc.lam_damp_iface, sc.lam_smooth_iface, 2.0, 8.0, 8.0,
             5.0, lx, lz, f"1 0.2 1.0 {sc.max_rel_pert}", 0]
    with open(os.path.join(wd, "invert2d.in"), "w") as f:
        f.write("\n".join(str(v) for v in lines) + "\n")
    out = run([INVERT2D], wd, "invert2d.log")
    print(f"[5] invert2d_hip: {sum(1 for ln in out.splitlines() if ln.startswith('Iteration'))} iterations")
    return out


def report(name: str, wd: str) -> dict:
    """RMS by iteration (the program's log), the recovery of the velocity
    perturbation (correlation and amplitude ratio of recovered vs true
    relative perturbation over the nodes below the seafloor whose ray density
    is non-zero) and, for the joint scenarios, the interface misfit before
    and after."""
    sc = SCENARIOS[name]
    log = os.path.join(wd, "invert_log_latest.txt")
    rms = []
    if os.path.exists(log):
        # "Iteration   k ndata=  n RMS=  r RMS_new=  r' ..." per iteration
        for ln in open(log):
            if ln.startswith("Iteration") and "RMS=" in ln:
                p = ln.replace("=", " ").split()
                rms.append(float(p[p.index("RMS") + 1]))
                last_new = float(p[p.index("RMS_new") + 1])
        if rms:
            rms.append(last_new)
    true = formats.read_smesh(os.path.join(wd, "true_vp.smesh"))
    init = formats.read_smesh(os.path.join(wd, "init_vp.smesh"))
    final_path = os.path.join(wd, "vgrid_1_vtype1.smesh")
    out = {"scenario": name, "rms_by_iteration": rms}
    if os.path.exists(final_path):
        fin = formats.read_smesh(final_path)
        dv_true = true.vel / init.vel - 1.0
        dv_rec = fin.vel / init.vel - 1.0
        # nodes below the seafloor that the inversion moved (ray coverage)
        mask = (np.abs(dv_rec) > 1e-6) & (init.zrel[None, :] > 0.0)
        a, b = dv_true[mask], dv_rec[mask]
        if a.size > 10:
            out["velocity_recovery"] = {"nodes": int(a.size), "correlation": float(np.corrcoef(a, b)[0, 1]),
                                        "amplitude_ratio": float(np.std(b) / np.std(a))}
    for k in sc.ifaces:
        fn = os.path.join(wd, f"iface_{k}.refl")
        if not os.path.exists(fn):
            continue
        tag = "moho" if k == 4 else "sed_bottom"
        xt, zt = formats.read_refl(os.path.join(wd, f"true_{tag}.dat"))
        xi, zi = formats.read_refl(os.path.join(wd, f"init_{tag}.dat"))
        xf, zf = formats.read_refl(fn)
        zf_on = np.interp(xt, xf, zf)
        zi_on = np.interp(xt, xi, zi)
        out[f"interface_{k}_rms_km"] = {"start": float(np.sqrt(np.mean((zi_on - zt) ** 2))),
                                        "final": float(np.sqrt(np.mean((zf_on - zt) ** 2)))}
    print("[report]", out)
    plot(name, wd, out)
    return out


def plot(name: str, wd: str, rep: dict) -> None:
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
    except ImportError:
        return
    fig, ax = plt.subplots(figsize=(6, 4))
    ax.plot(range(len(rep["rms_by_iteration"])), rep["rms_by_iteration"], "o-")
    ax.set_xlabel("iteration")
    ax.set_ylabel("RMS (s)")
    ax.set_title(f"checkerboard: {name}")
    fig.savefig(os.path.join(wd, "rms.png"), dpi=120, bbox_inches="tight")
    plt.close(fig)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("scenario", choices=sorted(SCENARIOS))
    ap.add_argument("--workdir", default=None, help="output folder (default ./example_<scenario>)")
    ap.add_argument("--dx", type=float, default=0.5, help="grid spacing, km")
    ap.add_argument("--niter", type=int, default=0, help="inversion iterations (default: the scenario's)")
    ap.add_argument("--seed", type=int, default=1, help="pick noise seed")
    ap.add_argument("--steps", default="all", help="comma list of prepare,forward,picks,invert,report or 'all'")
    a = ap.parse_args(argv)
    wd = os.path.abspath(a.workdir or f"example_{a.scenario}")
    steps = ["prepare", "forward", "picks", "invert", "report"] if a.steps == "all" else a.steps.split(",")
    for exe in (FMT2D, INVERT2D):
        if ("forward" in steps or "invert" in steps) and not os.path.exists(exe):
            raise SystemExit(f"{exe} is not built: python -m fmt2d_amd.build")
    if "prepare" in steps:
        prepare(a.scenario, wd, a.dx)
    if "forward" in steps:
        forward(wd)
    if "picks" in steps:
        picks(a.scenario, wd, a.seed)
    if "invert" in steps:
        invert(a.scenario, wd, a.niter)
    if "report" in steps:
        return report(a.scenario, wd)
    return None


if __name__ == "__main__":
    main()
