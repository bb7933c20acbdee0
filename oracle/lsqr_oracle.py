"""ORACLE -- numpy restatement of the reference's LSQR on the augmented,
smoothed and damped system (test infrastructure; only tests/, smoke() and
bench.py's cpu_baseline use it).

Restated from the reference sources:
  run_lsqr                    src/inversion2d.f90:6578-7167 (smooth_type 1:
                              [G; w L; wd I]; 2: [G; w R_h; w R_v; wd I]
                              with the interface R_if in the R_h block;
                              3: [G; w L_vel + w_if R_if; wd I]; jumping
                              right-hand side from dm_sum, column mask where
                              gtg_diag < 0.01 eps_min, atol 1e-4 on test2)
  apply_laplacian_smooth_2d   src/inversion2d.f90:1161-1249
  build_gauss_smooth_matrix,  src/inversion2d.f90:1252-1402 (R_h, R_v with
  apply_gauss_smooth_*        mvscale baked in; products :1405-1521)
  build_iface_gauss_smooth_matrix  src/inversion2d.f90:1524-1636
  sparse_Ax / sparse_ATx      src/inversion2d.f90:2997-3042
  gtg_diag / eps_min          src/inversion2d.f90:451-486
The reference's FP64 summation orders are kept: sequential row sums of G and
of its transpose (scipy's CSR kernels), the Laplacian's neighbour order, and
front-to-back dot products.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp


def read_lsqr_dump(path: str):
    """The first run_lsqr call of the reference checker build (oracle/raydump.inc)."""
    with open(path, "rb") as f:
        hi = np.fromfile(f, dtype=np.int32, count=7)
        ndata, nparm, nnz, n_vel, stype, maxit, iters = (int(x) for x in hi)
        hd = np.fromfile(f, dtype=np.float64, count=8)
        rowptr = np.fromfile(f, dtype=np.int32, count=ndata + 1)
        gindx = np.fromfile(f, dtype=np.int32, count=nnz)
        gval = np.fromfile(f, dtype=np.float64, count=nnz)
        res = np.fromfile(f, dtype=np.float64, count=ndata)
        gtg = np.fromfile(f, dtype=np.float64, count=nparm)
        dm_sum = np.fromfile(f, dtype=np.float64, count=nparm)
        dm = np.fromfile(f, dtype=np.float64, count=nparm)
        D = dict(ndata=ndata, nparm=nparm, nnz=nnz, n_vel=n_vel, smooth_type=stype, max_iter=maxit, iters=iters,
                 wdv=hd[0], wdv_iface=hd[1], w_smooth=hd[2], w_smooth_iface=hd[3], eps_min=hd[4], rnorm=hd[5],
                 bnorm=hd[6], test2=hd[7], rowptr=rowptr, gindx=gindx, gval=gval, residual=res, gtg=gtg,
                 dm_sum=dm_sum, dm=dm)
        # parameterisation and smoothing operators (later dumps only)
        mv = np.fromfile(f, dtype=np.float64, count=nparm)
        if len(mv) < nparm:
            return D
        D["mvscale"] = mv
        D["corr_x"], D["corr_z"] = (float(x) for x in np.fromfile(f, dtype=np.float64, count=2))
        nvg = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        if nvg > 0:
            vnx, vnz = (int(x) for x in np.fromfile(f, dtype=np.int32, count=2))
            D["vx"] = np.fromfile(f, dtype=np.float64, count=vnx)
            D["vz"] = np.fromfile(f, dtype=np.float64, count=vnz)
            # Fortran inv_idx(nx, nz): ix fastest -> (nz, nx)
            D["inv_idx"] = np.fromfile(f, dtype=np.int32, count=vnx * vnz).reshape(vnz, vnx)
        nif = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        D["ifaces"] = []
        for _ in range(nif):
            iid, nxi, start = (int(x) for x in np.fromfile(f, dtype=np.int32, count=3))
            D["ifaces"].append(dict(id=iid, start=start, x=np.fromfile(f, dtype=np.float64, count=nxi)))
        gflag, iflag = (int(x) for x in np.fromfile(f, dtype=np.int32, count=2))
        if gflag:
            hn, vn = (int(x) for x in np.fromfile(f, dtype=np.int32, count=2))
            D["rh"] = (np.fromfile(f, dtype=np.int32, count=nparm + 1), np.fromfile(f, dtype=np.int32, count=hn))
            D["rv"] = (np.fromfile(f, dtype=np.int32, count=nparm + 1), np.fromfile(f, dtype=np.int32, count=vn))
            D["rh"] += (np.fromfile(f, dtype=np.float64, count=hn),)
            D["rv"] += (np.fromfile(f, dtype=np.float64, count=vn),)
        if iflag:
            inn = int(np.fromfile(f, dtype=np.int32, count=1)[0])
            D["rif_nnz"] = inn
            if inn > 0:
                rp = np.fromfile(f, dtype=np.int32, count=nparm + 1)
                ci = np.fromfile(f, dtype=np.int32, count=inn)
                D["rif"] = (rp, ci, np.fromfile(f, dtype=np.float64, count=inn))
    return D


class Laplacian:
    """apply_laplacian_smooth_2d: 5-point stencil over the inverted velocity
    nodes, neighbours added in the reference's order (x-1, x+1, z-1, z+1) and
    the degree times the centre subtracted last; 1-D second difference along
    each inverted interface with one-sided ends."""

    def __init__(self, inv_idx: np.ndarray, nparm: int, ifaces=()):
        # inv_idx: (nz, nx) 1-based parameter numbers, 0 = not inverted
        self.nparm = nparm
        nz, nx = inv_idx.shape
        pad = np.zeros((nz + 2, nx + 2), dtype=np.int64)
        pad[1:-1, 1:-1] = inv_idx
        act = inv_idx > 0
        self.idx = inv_idx[act] - 1
        iz, ix = np.nonzero(act)
        # the reference visits iz outer, ix inner: np.nonzero's order
        self.nb = []
        for (dz, dx) in ((0, -1), (0, 1), (-1, 0), (1, 0)):
            j = pad[iz + 1 + dz, ix + 1 + dx]
            self.nb.append(j)
        self.deg = sum((j > 0).astype(np.float64) for j in self.nb)
        self.ifaces = list(ifaces)

    def __call__(self, vin):
        vout = np.zeros(self.nparm)
        acc = np.zeros(len(self.idx))
        for j in self.nb:
            ok = j > 0
            acc[ok] = acc[ok] + vin[j[ok] - 1]
        acc = acc - self.deg * vin[self.idx]
        vout[self.idx] = acc
        for (start, nxi) in self.ifaces:
            for ix in range(1, nxi + 1):
                idx = start + ix
                if idx < 1 or idx > self.nparm:
                    continue
                if ix == 1:
                    if idx + 1 <= self.nparm:
                        vout[idx - 1] = vout[idx - 1] + vin[idx] - vin[idx - 1]
                elif ix == nxi:
                    vout[idx - 1] = vout[idx - 1] + vin[idx - 2] - vin[idx - 1]
                else:
                    vout[idx - 1] = vout[idx - 1] + vin[idx - 2]
                    vout[idx - 1] = vout[idx - 1] + vin[idx]
                    vout[idx - 1] = vout[idx - 1] - 2.0 * vin[idx - 1]
        return vout


def _csr(rp, ci, val, n):
    """CSR from 1-based row pointers / columns, entries kept in their order."""
    return sp.csr_matrix((val, ci - 1, rp - 1), shape=(n, n))


def gauss_matrices(vx, vz, inv_idx, corr_x, corr_z, mvscale):
    """build_gauss_smooth_matrix: R_h (along x) and R_v (along z) over the
    inverted velocity nodes.  Row idx holds the neighbours within the
    window (x: |k| <= int(3 corr_x/dx) + 1 columns; z: |dz| <= 3 corr_z), in
    ascending order, each exp(-d^2/corr^2) * mvscale(neighbour), then the
    diagonal -sum(beta) * mvscale(idx).  Returns 1-based (rowptr, col, val)
    per matrix, as the reference stores them."""
    nz, nx = inv_idx.shape
    nparm = len(mvscale)
    dx = vx[1] - vx[0]
    kmax = int(3.0 * corr_x / dx) + 1
    kmin = -int(3.0 * corr_x / dx) - 1
    rows_h = [None] * nparm
    rows_v = [None] * nparm
    for iz in range(nz):
        for ix in range(nx):
            idx = int(inv_idx[iz, ix])
            if idx <= 0:
                continue
            cols, vals, bsum = [], [], 0.0
            for jx in range(max(1, ix + 1 + kmin), min(nx, ix + 1 + kmax) + 1):
                if jx == ix + 1:
                    continue
                jdx = int(inv_idx[iz, jx - 1])
                if jdx <= 0:
                    continue
                d = vx[jx - 1] - vx[ix]
                beta = math.exp(-(d * d) / (corr_x * corr_x))
                bsum = bsum + beta
                cols.append(jdx)
                vals.append(beta * mvscale[jdx - 1])
            cols.append(idx)
            vals.append(-bsum * mvscale[idx - 1])
            rows_h[idx - 1] = (cols, vals)
            cols, vals, bsum = [], [], 0.0
            for jz in range(1, nz + 1):
                if jz == iz + 1:
                    continue
                d = vz[jz - 1] - vz[iz]
                if abs(d) > 3.0 * corr_z:
                    continue
                jdx = int(inv_idx[jz - 1, ix])
                if jdx <= 0:
                    continue
                beta = math.exp(-(d * d) / (corr_z * corr_z))
                bsum = bsum + beta
                cols.append(jdx)
                vals.append(beta * mvscale[jdx - 1])
            cols.append(idx)
            vals.append(-bsum * mvscale[idx - 1])
            rows_v[idx - 1] = (cols, vals)
    return _pack_rows(rows_h, nparm), _pack_rows(rows_v, nparm)


def iface_gauss_matrix(ifaces, corr_x_iface, mvscale):
    """build_iface_gauss_smooth_matrix: 1-D Gaussian rows for the interface
    parameters start+1 .. start+nx of each inverted interface, same layout as
    R_h (neighbours ascending, diagonal last)."""
    nparm = len(mvscale)
    rows = [None] * nparm
    for itf in ifaces:
        x, start = itf["x"], itf["start"]
        nxi = len(x)
        dx = x[1] - x[0]
        kmax = int(3.0 * corr_x_iface / dx) + 1
        kmin = -int(3.0 * corr_x_iface / dx) - 1
        for ix in range(1, nxi + 1):
            idx = start + ix
            if idx < 1 or idx > nparm:
                continue
            cols, vals, bsum = [], [], 0.0
            for jx in range(max(1, ix + kmin), min(nxi, ix + kmax) + 1):
                if jx == ix:
                    continue
                jdx = start + jx
                if jdx < 1 or jdx > nparm:
                    continue
                d = x[jx - 1] - x[ix - 1]
                beta = math.exp(-(d * d) / (corr_x_iface * corr_x_iface))
                bsum = bsum + beta
                cols.append(jdx)
                vals.append(beta * mvscale[jdx - 1])
            cols.append(idx)
            vals.append(-bsum * mvscale[idx - 1])
            rows[idx - 1] = (cols, vals)
    return _pack_rows(rows, nparm)


def _pack_rows(rows, nparm):
    rp = np.ones(nparm + 1, dtype=np.int64)
    cols, vals = [], []
    for i, r in enumerate(rows):
        n = 0 if r is None else len(r[0])
        rp[i + 1] = rp[i] + n
        if r is not None:
            cols.extend(r[0])
            vals.extend(r[1])
    return rp, np.asarray(cols, dtype=np.int64), np.asarray(vals, dtype=np.float64)


def seqdot(a, b):
    """Front-to-back FP64 sum of a*b (the reference's dot_product)."""
    if len(a) == 0:
        return 0.0
    return float(np.cumsum(a * b)[-1])


def run_lsqr(G: sp.csr_matrix, residual, gtg, eps_min, n_vel, wdv, wdv_iface, w_smooth, w_smooth_iface, dm_sum,
             lap: Laplacian, max_iter=400, atol=1.0e-4, smooth_type=1, rh=None, rv=None, rif=None):
    """run_lsqr.  smooth_type 1: Laplacian block; 3: Laplacian on the
    velocity rows and R_if on the interface rows of that block; 2: blocks
    w R_h and w R_v (R_if in the R_h block's interface rows).  rh, rv, rif:
    1-based (rowptr, col, val) as built by gauss_matrices /
    iface_gauss_matrix (rif None: not built).  Returns dm, iterations,
    rnorm, bnorm, test2."""
    ndata, nparm = G.shape
    j = np.arange(nparm)
    vel = j < n_vel
    iface = ~vel
    masked = vel & (gtg < eps_min * 0.01)
    ws = np.where(vel, w_smooth, w_smooth_iface)
    wd = np.where(vel, wdv, wdv_iface)
    damp_on = wdv > 0.0 or wdv_iface > 0.0
    GT = G.T.tocsr()
    nblk = 2 if smooth_type == 2 else 1
    zero = np.zeros(nparm)
    RH = _csr(*rh, nparm) if rh is not None else None
    RV = _csr(*rv, nparm) if rv is not None else None
    RIF = _csr(*rif, nparm) if rif is not None else None
    RHT = RH.T.tocsr() if RH is not None else None
    RVT = RV.T.tocsr() if RV is not None else None

    def rif_apply(x):  # apply_iface_gauss_smooth_h: zero when not built
        return RIF @ x if RIF is not None else np.zeros(nparm)

    def smooth13(x):
        s = lap(x)
        out = np.where(masked, 0.0, ws * s)
        if smooth_type == 3:
            out[iface] = (w_smooth_iface * rif_apply(x))[iface] if w_smooth_iface > 0.0 else 0.0
        return out

    sm_on = (w_smooth > 0.0 or w_smooth_iface > 0.0) if smooth_type != 2 else (w_smooth > 0.0)
    gsi = smooth_type == 2 and w_smooth_iface > 0.0 and RIF is not None

    def blocks_fwd(x, rhs):
        # the smoothing blocks of A x (rhs: of the jumping right-hand side,
        # negated, as the reference forms it)
        if smooth_type != 2:
            b = smooth13(x) if sm_on else zero.copy()
            return [-b if (rhs and sm_on) else b]
        if w_smooth > 0.0:
            h = RH @ x
            v = RV @ x
            h[masked] = 0.0
            v[masked] = 0.0
            H = -(w_smooth * h) if rhs else w_smooth * h
            V = -(w_smooth * v) if rhs else w_smooth * v
        else:
            H, V = zero.copy(), zero.copy()
        if gsi:
            r = rif_apply(x)
            H[iface] = (-(w_smooth_iface * r) if rhs else w_smooth_iface * r)[iface]
        elif not rhs:
            H[iface] = 0.0
        return [H, V]

    def damp(x, rhs):
        if not damp_on:
            return zero.copy()
        return np.where(masked, 0.0, -wd * x if rhs else wd * x)

    def A(x):
        return np.concatenate([G @ x] + blocks_fwd(x, False) + [damp(x, False)])

    def AT(u):
        v = GT @ u[:ndata]
        if smooth_type != 2:
            if sm_on:
                v = v + smooth13(u[ndata:ndata + nparm])
        else:
            uh, uv = u[ndata:ndata + nparm], u[ndata + nparm:ndata + 2 * nparm]
            if w_smooth > 0.0:
                h = RHT @ uh
                vv = RVT @ uv
                h[masked] = 0.0
                vv[masked] = 0.0
                v = v + w_smooth * h + w_smooth * vv
            if gsi:
                v[iface] = v[iface] + (w_smooth_iface * rif_apply(uh))[iface]
        if damp_on:
            ud = u[ndata + nblk * nparm:]
            v = v + np.where(masked, 0.0, wd * ud)
        return v

    u = np.concatenate([residual] + blocks_fwd(dm_sum, True) + [damp(dm_sum, True)])
    bnorm = float(np.sqrt(seqdot(u, u)))
    dm = np.zeros(nparm)
    if bnorm < 1e-30:
        return dm, 0, 0.0, bnorm, 0.0
    beta = bnorm
    u = u / beta
    v = AT(u)
    v[masked] = 0.0
    alpha = float(np.sqrt(seqdot(v, v)))
    if alpha < 1e-30:
        return dm, 0, bnorm, bnorm, 0.0
    v = v / alpha
    w = v.copy()
    phibar, rhobar = beta, alpha
    anorm2 = alpha ** 2 + beta ** 2
    it = 0
    test2 = 0.0
    rnorm = 0.0
    for it in range(1, max_iter + 1):
        v[masked] = 0.0
        u = A(v) - alpha * u
        beta = float(np.sqrt(seqdot(u, u)))
        if beta > 0.0:
            u = u / beta
        v = AT(u) - beta * v
        v[masked] = 0.0
        alpha = float(np.sqrt(seqdot(v, v)))
        if alpha > 0.0:
            v = v / alpha
        anorm2 = anorm2 + alpha ** 2 + beta ** 2
        rho = np.sqrt(rhobar ** 2 + beta ** 2)
        c = rhobar / rho
        s = beta / rho
        theta = s * alpha
        rhobar = -c * alpha
        phi = c * phibar
        phibar = s * phibar
        z = np.where(masked, 0.0, w)
        w = np.where(masked, 0.0, v - (theta / rho) * w)
        dm = dm + (phi / rho) * z
        rnorm = abs(phibar)
        arnorm = phibar * alpha * abs(c)
        test2 = arnorm / (np.sqrt(anorm2) * rnorm + 1.0e-30)
        if test2 < atol or 1.0 + test2 <= 1.0:
            break
    return dm, it, rnorm, bnorm, test2


def csr_from_dump(D):
    rp = D["rowptr"] - 1
    return sp.csr_matrix((D["gval"], D["gindx"] - 1, rp), shape=(D["ndata"], D["nparm"]))
