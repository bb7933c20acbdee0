"""Readers for the reference's text input formats.

* ``read_smesh``: ``nx nz vwater vair``, then ``xpos(nx)``, ``topo(nx)``,
  ``zpos_rel(nz)``, then ``vel(ix,1:nz)`` for ix=1..nx, whitespace-token
  based (``src/read_tomo2d.f90:8-79``).
* ``read_refl``: ``x z`` per non-empty line, malformed lines skipped
  (``src/read_tomo2d.f90:83-137``).
* ``read_geom``: ``nobs``; per OBS ``S x z nrec``; then ``nrec`` lines
  ``R x z ray_type t_obs unc`` (``src/read_tomo2d.f90:141-248``).
* ``read_fmt2d_in``: the 8-line ``fmt2d.in`` (``src/fmt2d_main.f90:59-68``).
"""
from __future__ import annotations

import dataclasses
from typing import List

import numpy as np


@dataclasses.dataclass
class Smesh:
    xpos: np.ndarray
    topo: np.ndarray
    zrel: np.ndarray
    vel: np.ndarray  # (nx, nz), iz fastest
    vwater: float
    vair: float


@dataclasses.dataclass
class Geom:
    obs_x: np.ndarray
    obs_z: np.ndarray
    rec_x: np.ndarray
    rec_z: np.ndarray
    obs_id: np.ndarray      # 1-based OBS index per receiver line
    ray_type: np.ndarray
    t_obs: np.ndarray
    unc: np.ndarray


def read_smesh(path: str) -> Smesh:
    with open(path) as f:
        tok = f.read().split()
    nx, nz = int(tok[0]), int(tok[1])
    if nx < 2 or nz < 2:
        raise ValueError(f"smesh grid too small: nx={nx} nz={nz}")
    vwater, vair = float(tok[2]), float(tok[3])
    vals = np.array(tok[4:4 + 2 * nx + nz + nx * nz], dtype=float)
    if vals.size != 2 * nx + nz + nx * nz:
        raise ValueError("smesh file truncated")
    xpos = vals[:nx]
    topo = vals[nx:2 * nx]
    zrel = vals[2 * nx:2 * nx + nz]
    vel = vals[2 * nx + nz:].reshape(nx, nz)
    return Smesh(xpos, topo, zrel, vel, vwater, vair)


def read_refl(path: str):
    xs: List[float] = []
    zs: List[float] = []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            try:
                x, z = float(p[0]), float(p[1])
            except (ValueError, IndexError):
                continue
            xs.append(x)
            zs.append(z)
    if not xs:
        raise ValueError(f"refl file has no valid data: {path}")
    return np.array(xs), np.array(zs)


def read_geom(path: str) -> Geom:
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    nobs = int(lines[0][0])
    if nobs < 1:
        raise ValueError(f"invalid nobs from geom file: {nobs}")
    ox, oz = [], []
    rx, rz, oid, rt, tt, uu = [], [], [], [], [], []
    i = 1
    for iobs in range(1, nobs + 1):
        h = lines[i]
        i += 1
        ox.append(float(h[1]))
        oz.append(float(h[2]))
        for _ in range(int(h[3])):
            r = lines[i]
            i += 1
            rx.append(float(r[1]))
            rz.append(float(r[2]))
            rt.append(int(r[3]))
            tt.append(float(r[4]))
            uu.append(float(r[5]))
            oid.append(iobs)
    return Geom(np.array(ox), np.array(oz), np.array(rx), np.array(rz),
                np.array(oid, dtype=np.int32), np.array(rt, dtype=np.int32),
                np.array(tt), np.array(uu))


def read_fmt2d_in(path: str) -> dict:
    with open(path) as f:
        ln = [l.rstrip("\n") for l in f]
    return dict(smesh=ln[0].strip(), refl=ln[1].strip(), sed=ln[2].strip(),
                geom=ln[3].strip(), write_rays=ln[4].strip().upper().startswith(("T", ".T")),
                arrivals=ln[5].strip(), obs=ln[6].strip(), rays_prefix=ln[7].strip())
