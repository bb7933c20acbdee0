"""ctypes wrapper of the C oracle (oracle/fmm2d_oracle.c).

ORACLE / TEST INFRASTRUCTURE: imported only by tests/, the smoke check and
bench.py's cpu_baseline leg.  Never used by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_fmt2d.so")
TF_REGION = [0, 5, 2, 2, 6, 6, 1, 1, 1, 1, 1, 5, 1]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _lib():
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    lib.orc_model_create.restype = C.c_void_p
    lib.orc_model_create.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double,
                                     C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    for name in ("orc_region_nnode", "orc_isec_nnode"):
        getattr(lib, name).argtypes = [C.c_void_p, C.c_int]
        getattr(lib, name).restype = C.c_int
    for name in ("orc_grid_nx", "orc_grid_nz"):
        getattr(lib, name).argtypes = [C.c_void_p]
        getattr(lib, name).restype = C.c_int
    lib.orc_forward_source.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int,
                                       C.POINTER(_dp), C.POINTER(_dp), C.POINTER(_dp)]
    lib.orc_region_nodes.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _dp, _dp]
    lib.orc_region_rnode.argtypes = [C.c_void_p, C.c_int, _ip]
    lib.orc_fully_regular.argtypes = [C.c_void_p, _ip]
    lib.orc_node_region.argtypes = [C.c_void_p, _ip]
    lib.orc_isec.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.orc_iface_z_at.argtypes = [C.c_void_p, C.c_int, C.c_double]
    lib.orc_iface_z_at.restype = C.c_double
    return lib


_L = None


def lib():
    global _L
    if _L is None:
        _L = _lib()
    return _L


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class OracleModel:
    """The reference model build (fmt2d_main.f90:91-418) restated in C."""

    def __init__(self, smesh, sed_x, sed_z, moho_x, moho_z):
        L = lib()
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (smesh.xpos, smesh.topo, smesh.zrel, smesh.vel, sed_x, sed_z, moho_x, moho_z)]
        xp, tp, zr, vl, sx, sz, mx, mz = self._keep
        self.h = L.orc_model_create(len(xp), len(zr), _p(xp), _p(tp), _p(zr), _p(vl),
                                    float(smesh.vwater), len(sx), _p(sx), _p(sz), len(mx), _p(mx), _p(mz))
        self.nx = L.orc_grid_nx(self.h)
        self.nz = L.orc_grid_nz(self.h)

    def region_nnode(self, r):
        return lib().orc_region_nnode(self.h, r)

    def region_nodes(self, r):
        n = self.region_nnode(r)
        i1, i2, i3 = (np.zeros(n, np.int32) for _ in range(3))
        x, z, v = (np.zeros(n) for _ in range(3))
        lib().orc_region_nodes(self.h, r, _p(i1, _ip), _p(i2, _ip), _p(i3, _ip), _p(x), _p(z), _p(v))
        return dict(i1=i1, i2=i2, i3=i3, x=x, z=z, vel=v)

    def region_rnode(self, r):
        out = np.zeros(self.nx * self.nz, np.int32)
        lib().orc_region_rnode(self.h, r, _p(out, _ip))
        return out.reshape(self.nz, self.nx)

    def fully_regular(self):
        out = np.zeros(self.nx * self.nz, np.int32)
        lib().orc_fully_regular(self.h, _p(out, _ip))
        return out.reshape(self.nz, self.nx)

    def node_region(self):
        out = np.zeros(self.nx * self.nz, np.int32)
        lib().orc_node_region(self.h, _p(out, _ip))
        return out.reshape(self.nz, self.nx)

    def isec(self, k):
        n = lib().orc_isec_nnode(self.h, k)
        arrs = [np.zeros(n) for _ in range(6)]
        lib().orc_isec(self.h, k, *[_p(a) for a in arrs])
        return dict(zip(("x", "z", "nx", "nz", "vel_top", "vel_bot"), arrs))

    def seafloor_z(self, x):
        return lib().orc_iface_z_at(self.h, 2, float(x))

    def forward_source(self, sx, nstage=12, headwave=True):
        """Stage fields of one source: list of 12 (T, gx, gz) arrays in the
        reference's region node order (None for skipped stages)."""
        outs = []
        PT, PX, PZ = (_dp * 12)(), (_dp * 12)(), (_dp * 12)()
        for t in range(1, 13):
            n = self.region_nnode(TF_REGION[t])
            a = [np.full(n, np.nan) for _ in range(3)]
            outs.append(a)
            PT[t - 1], PX[t - 1], PZ[t - 1] = _p(a[0]), _p(a[1]), _p(a[2])
        lib().orc_forward_source(self.h, float(sx), int(nstage), int(bool(headwave)), PT, PX, PZ)
        return [tuple(o) if t + 1 <= nstage else None for t, o in enumerate(outs)]
