"""Reader of the stream file written by oracle/ref_fields.f90 (the reference
modules' own model build and stage fields).  Test infrastructure only."""
from __future__ import annotations

import numpy as np

TF_REGION = [0, 5, 2, 2, 6, 6, 1, 1, 1, 1, 1, 5, 1]
# run_forward2d visiting order of the tf slots (tf1 is recomputed for path 4)
STAGE_ORDER = [1, 2, 3, 4, 5, 1, 6, 7, 8, 9, 10, 11, 12]


def read_refdump(path: str, nsrc: int, nstage: int = 12):
    buf = open(path, "rb").read()
    off = 0

    def ints(n):
        nonlocal off
        a = np.frombuffer(buf, dtype="<i4", count=n, offset=off)
        off += 4 * n
        return a.copy()

    def dbls(n):
        nonlocal off
        a = np.frombuffer(buf, dtype="<f8", count=n, offset=off)
        off += 8 * n
        return a.copy()

    nx, nz = ints(2)
    dx, x0, z0 = dbls(3)
    out = dict(nx=int(nx), nz=int(nz), dx=dx, x0=x0, z0=z0)
    out["fully_regular"] = ints(nx * nz).reshape(nz, nx)
    out["node_region"] = ints(nx * nz).reshape(nz, nx)
    regs = {}
    for r in range(1, 7):
        nnode, ngnode = ints(2)
        i1, i2, i3 = ints(nnode), ints(nnode), ints(nnode)
        x, z, v = dbls(nnode), dbls(nnode), dbls(nnode)
        rn = ints(nx * nz).reshape(nz, nx)
        regs[r] = dict(nnode=int(nnode), ngnode=int(ngnode), i1=i1, i2=i2, i3=i3, x=x, z=z,
                       vel=v, rnode=rn)
    out["regions"] = regs
    isecs = {}
    for k in range(1, 6):
        (n,) = ints(1)
        x, z = dbls(n), dbls(n)
        nrm = dbls(2 * n).reshape(n, 2)
        vt, vb = dbls(n), dbls(n)
        isecs[k] = dict(x=x, z=z, nx=nrm[:, 0], nz=nrm[:, 1], vel_top=vt, vel_bot=vb)
    out["isec"] = isecs
    srcs = []
    order = [t for t in STAGE_ORDER if t <= nstage]
    for _ in range(nsrc):
        st = []
        for _t in order:
            itf, rid = ints(2)
            nn = regs[int(rid)]["nnode"]
            T = dbls(nn)
            g = dbls(2 * nn).reshape(nn, 2)
            st.append((int(itf), T, g[:, 0].copy(), g[:, 1].copy()))
        srcs.append(st)
    out["sources"] = srcs
    assert off == len(buf), (off, len(buf))
    return out
