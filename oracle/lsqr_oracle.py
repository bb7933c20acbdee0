"""ORACLE -- numpy restatement of the reference's LSQR on the augmented,
smoothed and damped system (test infrastructure; only tests/, smoke() and
bench.py's cpu_baseline use it).

Restated from the reference sources:
  run_lsqr                    src/inversion2d.f90:6578-7167 (smooth_type 1:
                              [G; w L; wd I], jumping right-hand side
                              -w L dm_sum, -wd dm_sum, column mask where
                              gtg_diag < 0.01 eps_min, atol 1e-4 on test2)
  apply_laplacian_smooth_2d   src/inversion2d.f90:1161-1249
  sparse_Ax / sparse_ATx      src/inversion2d.f90:2997-3042
  gtg_diag / eps_min          src/inversion2d.f90:451-486
The reference's FP64 summation orders are kept: sequential row sums of G and
of its transpose (scipy's CSR kernels), the Laplacian's neighbour order, and
front-to-back dot products.
"""
from __future__ import annotations

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
    return dict(ndata=ndata, nparm=nparm, nnz=nnz, n_vel=n_vel, smooth_type=stype, max_iter=maxit, iters=iters,
                wdv=hd[0], wdv_iface=hd[1], w_smooth=hd[2], w_smooth_iface=hd[3], eps_min=hd[4], rnorm=hd[5],
                bnorm=hd[6], test2=hd[7], rowptr=rowptr, gindx=gindx, gval=gval, residual=res, gtg=gtg,
                dm_sum=dm_sum, dm=dm)


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


def seqdot(a, b):
    """Front-to-back FP64 sum of a*b (the reference's dot_product)."""
    if len(a) == 0:
        return 0.0
    return float(np.cumsum(a * b)[-1])


def run_lsqr(G: sp.csr_matrix, residual, gtg, eps_min, n_vel, wdv, wdv_iface, w_smooth, w_smooth_iface, dm_sum,
             lap: Laplacian, max_iter=400, atol=1.0e-4):
    """run_lsqr with smooth_type 1 (Laplacian).  Returns dm, iterations,
    rnorm, bnorm, test2."""
    ndata, nparm = G.shape
    j = np.arange(nparm)
    vel = j < n_vel
    masked = vel & (gtg < eps_min * 0.01)
    ws = np.where(vel, w_smooth, w_smooth_iface)
    wd = np.where(vel, wdv, wdv_iface)
    smooth_on = w_smooth > 0.0 or w_smooth_iface > 0.0
    damp_on = wdv > 0.0 or wdv_iface > 0.0
    GT = G.T.tocsr()

    def smooth(x):
        s = lap(x)
        return np.where(masked, 0.0, ws * s)

    def A(x):
        out = [G @ x]
        out.append(smooth(x) if smooth_on else np.zeros(nparm))
        out.append(np.where(masked, 0.0, wd * x) if damp_on else np.zeros(nparm))
        return np.concatenate(out)

    def AT(u):
        v = GT @ u[:ndata]
        if smooth_on:
            v = v + smooth(u[ndata:ndata + nparm])
        if damp_on:
            v = v + np.where(masked, 0.0, wd * u[ndata + nparm:])
        return v

    u = np.concatenate([residual,
                        -smooth(dm_sum) if smooth_on else np.zeros(nparm),
                        np.where(masked, 0.0, -wd * dm_sum) if damp_on else np.zeros(nparm)])
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
