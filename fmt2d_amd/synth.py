"""Deterministic synthetic marine models of the FMT2D model family.

The files written here follow the reference's text formats
(``src/read_tomo2d.f90``): ``.smesh`` (``:8-79``), interface ``x z`` lines
(``:83-137``) and ``geom`` (``:141-248``), so the same inputs can be fed to the
reference executables and to this package.

Model family (all depths in km below sea level, velocities in km/s):

* water 1.5 km/s above a seafloor ``3 + 1*sin(2*pi*x/80 + 0.1)``;
* sediments 1.8 -> 3.0 km/s, linear in depth below the seafloor, thickness
  ``1.5 + 0.5*sin(2*pi*x/60 + 0.7)`` (1-2 km);
* crust 4.5 -> 7.0 km/s, linear from the sediment base to the Moho;
* Moho ``15 + 2*sin(2*pi*x/70 + 0.1)``; mantle 8.0 km/s;
* optional checkerboard ``v *= 1 + a*sign(sin(pi*x/Lx)*sin(pi*z/Lz))`` below
  the seafloor;
* optional seeded random Fourier perturbation of the sediment base and Moho.

The ``.smesh`` depth axis is relative to the seafloor with the same spacing
as x; its last node is placed so that the reference's propagation grid
(``zmax = zrel(nz) + max(topo) + 5``, ``src/fmt2d_main.f90:200-205``) ends at
``z_bottom``.
"""
from __future__ import annotations

import dataclasses
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

VWATER = 1.5
VAIR = 0.33


@dataclasses.dataclass
class Model:
    xpos: np.ndarray          # (nx,)
    topo: np.ndarray          # (nx,) seafloor depth
    zrel: np.ndarray          # (nzr,) depth below seafloor of .smesh rows
    vel: np.ndarray           # (nx, nzr) smesh velocities (ix-major like the file)
    sed_z: np.ndarray         # (nx,) sediment-base depth
    moho_z: np.ndarray        # (nx,) Moho depth
    vwater: float = VWATER
    vair: float = VAIR


@dataclasses.dataclass
class Geometry:
    obs_x: np.ndarray                     # (nobs,)
    obs_z: np.ndarray                     # (nobs,)
    # per OBS: list of (x, z, ray_type, t_obs, unc)
    receivers: List[List[Tuple[float, float, int, float, float]]]

    @property
    def nrec(self) -> int:
        return sum(len(r) for r in self.receivers)


# Phases of the seafloor and Moho sinusoids keep interface nodes off the grid
# rows: an interface node lying exactly on a grid node makes the reference's
# interface re-seeding produce a 0/0 gradient and a heap-order tie.
SEAFLOOR_PHASE = 0.1
MOHO_PHASE = 0.1


def seafloor(x):
    return 3.0 + 1.0 * np.sin(2.0 * np.pi * x / 80.0 + SEAFLOOR_PHASE)


def sediment_thickness(x):
    return 1.5 + 0.5 * np.sin(2.0 * np.pi * x / 60.0 + 0.7)


def moho(x, wavelength=70.0):
    return 15.0 + 2.0 * np.sin(2.0 * np.pi * x / wavelength + MOHO_PHASE)


def fourier_perturbation(x, rng, nmodes=6, amp_total=1.5, lmin=10.0, lmax=100.0):
    """Sum of ``nmodes`` sines with wavelengths U[lmin,lmax] and sum|a_k|<=amp_total."""
    lam = rng.uniform(lmin, lmax, nmodes)
    phi = rng.uniform(0.0, 2.0 * np.pi, nmodes)
    a = rng.uniform(-1.0, 1.0, nmodes)
    a *= amp_total / max(np.sum(np.abs(a)), 1e-12)
    out = np.zeros_like(x)
    for k in range(nmodes):
        out += a[k] * np.sin(2.0 * np.pi * x / lam[k] + phi[k])
    return out


def make_model(dx: float = 0.5, xmax: float = 200.0, z_bottom: float = 30.0,
               checker_amp: float = 0.05, checker_lx: float = 20.0,
               checker_lz: float = 5.0, fourier_seed: Optional[int] = None) -> Model:
    nx = int(round(xmax / dx)) + 1
    xpos = np.round(np.arange(nx) * dx, 9)
    topo = seafloor(xpos)
    sed_z = topo + sediment_thickness(xpos)
    moho_z = moho(xpos)
    if fourier_seed is not None:
        rng = np.random.default_rng(fourier_seed)
        sed_z = sed_z + fourier_perturbation(xpos, rng, amp_total=0.5)
        moho_z = moho_z + fourier_perturbation(xpos, rng, amp_total=1.5)
        sed_z = np.maximum(sed_z, topo + 0.3)
        moho_z = np.maximum(moho_z, sed_z + 2.0)
    zrel_max = z_bottom - 5.0 - topo.max()
    nzr = int(np.floor(zrel_max / dx + 1e-9)) + 1
    zrel = np.round(np.arange(nzr) * dx, 9)
    zrel[-1] = zrel_max
    vel = np.zeros((nx, nzr))
    for ix in range(nx):
        z = topo[ix] + zrel
        hs = sed_z[ix] - topo[ix]
        v = np.where(zrel < hs, 1.8 + 1.2 * zrel / hs,
                     np.where(z < moho_z[ix],
                              4.5 + 2.5 * (z - sed_z[ix]) / (moho_z[ix] - sed_z[ix]),
                              8.0))
        if checker_amp:
            sgn = np.sign(np.sin(np.pi * xpos[ix] / checker_lx) * np.sin(np.pi * z / checker_lz))
            v = np.where(zrel > 0.0, v * (1.0 + checker_amp * sgn), v)
        vel[ix] = v
    return Model(xpos=xpos, topo=topo, zrel=zrel, vel=vel, sed_z=sed_z, moho_z=moho_z)


def make_geometry(obs_x: Sequence[float], shot_x: Sequence[float], model: Model,
                  shot_z: float = 0.010, ray_types: Sequence[int] = (1,),
                  max_offset: Optional[float] = None) -> Geometry:
    obs_x = np.asarray(obs_x, dtype=float)
    obs_z = np.interp(obs_x, model.xpos, model.topo)
    recs = []
    for ox in obs_x:
        lst = []
        for rt in ray_types:
            for sx in shot_x:
                if max_offset is not None and abs(sx - ox) > max_offset:
                    continue
                lst.append((float(sx), float(shot_z), int(rt), 0.0, 0.05))
        recs.append(lst)
    return Geometry(obs_x=obs_x, obs_z=obs_z, receivers=recs)


# ---------------------------------------------------------------- writers

def write_smesh(path: str, m: Model) -> None:
    nx, nz = m.vel.shape
    with open(path, "w") as f:
        f.write(f" {nx:d} {nz:d} {m.vwater:.6f} {m.vair:.6f}\n")
        f.write(" ".join(f"{v:.6f}" for v in m.xpos) + "\n")
        f.write(" ".join(f"{v:.10f}" for v in m.topo) + "\n")
        f.write(" ".join(f"{v:.10f}" for v in m.zrel) + "\n")
        for ix in range(nx):
            f.write(" ".join(f"{v:.10f}" for v in m.vel[ix]) + "\n")


def write_interface(path: str, x: np.ndarray, z: np.ndarray) -> None:
    with open(path, "w") as f:
        for xi, zi in zip(x, z):
            f.write(f"{xi:.6f} {zi:.10f}\n")


def write_geom(path: str, g: Geometry) -> None:
    with open(path, "w") as f:
        f.write(f"{len(g.obs_x)}\n")
        for i, (ox, oz) in enumerate(zip(g.obs_x, g.obs_z)):
            f.write(f"S {ox:.6f} {oz:.6f} {len(g.receivers[i])}\n")
            for (rx, rz, rt, t, u) in g.receivers[i]:
                f.write(f"R {rx:.6f} {rz:.6f} {rt:d} {t:.6f} {u:.6f}\n")


def write_fmt2d_case(outdir: str, m: Model, g: Geometry, write_rays: bool = False) -> None:
    """Write the model, geometry and an ``fmt2d.in`` (8 lines,
    ``src/fmt2d_main.f90:59-68``) into ``outdir``."""
    os.makedirs(outdir, exist_ok=True)
    write_smesh(os.path.join(outdir, "model.smesh"), m)
    write_interface(os.path.join(outdir, "moho.dat"), m.xpos, m.moho_z)
    write_interface(os.path.join(outdir, "sed.dat"), m.xpos, m.sed_z)
    write_geom(os.path.join(outdir, "geom.dat"), g)
    with open(os.path.join(outdir, "fmt2d.in"), "w") as f:
        f.write("model.smesh\nmoho.dat\nsed.dat\ngeom.dat\n")
        f.write("T\n" if write_rays else "F\n")
        f.write("arrivals.dat\nobs_times.dat\nrays_src\n")


# ---------------------------------------------------------------- named workloads

def plain_pg_case(dx: float = 0.5):
    """The CPU-runnable plain case: 200 km x 30 km at ``dx`` (401x61 at 0.5 km),
    20 OBS every 10 km, shots every 0.5 km at 10 m depth, Pg only."""
    m = make_model(dx=dx)
    obs = np.arange(5.0, 200.0, 10.0)
    shots = np.round(np.arange(0.5, 200.0 - 0.25, 0.5), 6)
    g = make_geometry(obs, shots, m, ray_types=(1,))
    return m, g


def multistage_case(dx: float = 0.1, nobs: int = 60):
    """The full multistage forward: 0.1 km grid (2001x301), ``nobs`` OBS every
    3.3 km, 1000 shots every 0.2 km at 10 m depth."""
    m = make_model(dx=dx)
    obs = np.round(2.0 + 3.3 * np.arange(nobs), 6)
    shots = np.round(0.1 + 0.2 * np.arange(1000), 6)
    g = make_geometry(obs, shots, m, ray_types=(1,))
    return m, g


def multistage_phases_case(dx: float = 0.1, obs_index: Sequence[int] = (0, 30, 59)):
    """The full multistage forward's model and shots (``multistage_case``)
    with receiver lines of all five phases in the reference's windows for the
    OBS ``obs_index`` of its 60 (x = 2.0 + 3.3 k km)."""
    m = make_model(dx=dx)
    obs = np.round(2.0 + 3.3 * np.asarray(obs_index, dtype=float), 6)
    shots = np.round(0.1 + 0.2 * np.arange(1000), 6)
    return m, phase_windows_geometry(m, obs, shots)


def batched_case(dx: float = 0.1, nobs: int = 200, seed: int = 3):
    """The batched 1000-source-stage forward: 0.1 km grid (2001x301), sediment
    base and Moho perturbed by seeded random Fourier modes (wavelengths
    10-100 km, sum of |amplitudes| 0.5 km and 1.5 km), ``nobs`` OBS every
    200/nobs km from 0.5 km (every 1 km for 200), 1000 shots every 0.2 km at
    10 m depth, receiver lines of all five phases in the reference's
    windows."""
    m = make_model(dx=dx, fourier_seed=seed)
    step = 200.0 / nobs
    obs = np.round(step / 2.0 + step * np.arange(nobs), 6)
    shots = np.round(0.1 + 0.2 * np.arange(1000), 6)
    return m, phase_windows_geometry(m, obs, shots)


def gradient_case(dx: float = 0.1, nobs: int = 60):
    """The adjoint-gradient workload: 0.1 km grid (2001x301), ``nobs`` OBS
    every 3.3 km, 1000 shots every 0.2 km at 10 m depth, Pg at every offset
    and PmP at 3-30 km; the true model carries a +-5 % checkerboard of
    10 km x 3 km cells, the starting model none.
    Returns (true model, starting model, geometry without picks)."""
    m_true = make_model(dx=dx, checker_lx=10.0, checker_lz=3.0)
    m0 = make_model(dx=dx, checker_amp=0.0)
    obs = np.round(2.0 + 3.3 * np.arange(nobs), 6)
    shots = np.round(0.1 + 0.2 * np.arange(1000), 6)
    recs = []
    for ox in obs:
        lst = [(float(sx), 0.010, 1, 0.0, 0.03) for sx in shots]
        lst += [(float(sx), 0.010, 3, 0.0, 0.03) for sx in shots if 3.0 <= abs(sx - ox) <= 30.0]
        recs.append(lst)
    g = Geometry(obs_x=obs, obs_z=np.interp(obs, m_true.xpos, m_true.topo), receivers=recs)
    return m_true, m0, g


def phase_windows_geometry(model: Model, obs_x: Sequence[float], shot_x: Sequence[float],
                           shot_z: float = 0.010) -> Geometry:
    """Receiver lines for all five phases with the reference's windows:
    Pg at every offset, PsedP at 3-6 mean sediment thicknesses, PmP at 3-30 km,
    Pg_ss and Pg_ss2 up to 50 km (ray types 1..5, ``fmt2d_main.f90:686-735``)."""
    obs_x = np.asarray(obs_x, dtype=float)
    hsed = float(np.mean(model.sed_z - model.topo))
    recs = []
    for ox in obs_x:
        lst = []
        for sx in shot_x:
            off = abs(sx - ox)
            lst.append((float(sx), shot_z, 1, 1.0, 0.05))
            if 3.0 * hsed <= off <= 6.0 * hsed:
                lst.append((float(sx), shot_z, 2, 1.0, 0.05))
            if 3.0 <= off <= 30.0:
                lst.append((float(sx), shot_z, 3, 1.0, 0.05))
            if off <= 50.0:
                lst.append((float(sx), shot_z, 4, 1.0, 0.05))
                lst.append((float(sx), shot_z, 5, 1.0, 0.05))
        recs.append(lst)
    return Geometry(obs_x=obs_x, obs_z=np.interp(obs_x, model.xpos, model.topo), receivers=recs)


def multiphase_case(dx: float = 0.5):
    """The plain case's model and sources with receiver lines for all five
    phases (see ``phase_windows_geometry``)."""
    m = make_model(dx=dx)
    obs = np.arange(5.0, 200.0, 10.0)
    shots = np.round(np.arange(0.5, 200.0 - 0.25, 0.5), 6)
    return m, phase_windows_geometry(m, obs, shots)


def inversion_case(dx: float = 0.5, obs_x: Sequence[float] = (45.0, 95.0, 145.0), shot_dx: float = 1.0,
                   phases: bool = False):
    """A small inversion case: the plain model family without the
    checkerboard as the starting model, the checkerboard model as the truth,
    ``obs_x`` OBS and shots every ``shot_dx`` km at 10 m depth, Pg only (or,
    with ``phases``, all five phases in the reference's windows).
    Returns (true model, starting model, geometry without picks)."""
    m_true = make_model(dx=dx)
    m0 = make_model(dx=dx, checker_amp=0.0)
    shots = np.round(np.arange(shot_dx / 2.0, 200.0 - shot_dx / 4.0, shot_dx), 6)
    if phases:
        g = phase_windows_geometry(m_true, obs_x, shots)
    else:
        g = make_geometry(obs_x, shots, m_true, ray_types=(1,))
    return m_true, m0, g


def with_picks(g: Geometry, t_calc, valid, sigma: float = 0.05, seed: int = 1) -> Geometry:
    """Geometry whose receiver lines carry picks t_calc + N(0, sigma) (seeded,
    drawn line by line) and -1 where the ray is invalid."""
    rng = np.random.default_rng(seed)
    k = 0
    recs = []
    for rl in g.receivers:
        lst = []
        for (x, z, rt, t, u) in rl:
            noise = rng.normal(0.0, sigma)
            tobs = float(t_calc[k]) + noise if valid[k] else -1.0
            lst.append((x, z, rt, tobs, sigma))
            k += 1
        recs.append(lst)
    return Geometry(obs_x=g.obs_x, obs_z=g.obs_z, receivers=recs)
