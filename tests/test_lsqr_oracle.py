"""Pins of the LSQR oracle (oracle/lsqr_oracle.py), no GPU needed.

* against the reference: the first run_lsqr solve of the reference
  inversion (tests/golden/ref_lsqr_small.npz, from the reference built from
  its sources) is reproduced bit for bit from its own system;
* against the textbook: with the stopping tolerance tightened, the
  recurrences converge to scipy's LSQR solution of the same explicit
  augmented system [G; w L; wd I] (no masked columns, so the operator is
  exactly that matrix);
* the Laplacian block against its closed-form stencil on a small grid.
"""
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle.lsqr_oracle import Laplacian, run_lsqr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "ref_lsqr_small.npz")


def inv_grid(nx=401, nz=61, dx=0.5):
    from fmt2d_amd import synth
    m = synth.make_model(dx=dx, checker_amp=0.0)
    inv = np.zeros((nz, nx), dtype=np.int64)
    n = 0
    for iz in range(nz):
        for ix in range(nx):
            ivx = min(ix + 1, nx - 1)
            wx = (ix + 1) - ivx
            zsf = (1.0 - wx) * m.topo[ivx - 1] + wx * m.topo[ivx]
            if iz * dx - zsf >= 0.0:
                n += 1
                inv[iz, ix] = n
    return inv, n


def test_oracle_reproduces_reference_lsqr():
    Z = np.load(FIX)
    ndata, nparm, nnz, n_vel, stype, maxit, iters = (int(x) for x in Z["scal"])
    wdv, wdv_i, ws, ws_i, eps, rnorm, bnorm, test2 = Z["par"]
    inv, n = inv_grid()
    assert n == n_vel
    G = sp.csr_matrix((Z["gval"], Z["gindx"] - 1, Z["rowptr"] - 1), shape=(ndata, nparm))
    dm, it, rn, bn, t2 = run_lsqr(G, Z["residual"], Z["gtg"], eps, n_vel, wdv, wdv_i, ws, ws_i, Z["dm_sum"],
                                  Laplacian(inv, nparm), maxit)
    assert it == iters
    assert rn == rnorm and bn == bnorm and t2 == test2
    assert np.array_equal(dm, Z["dm"])


def test_recurrences_converge_to_textbook_lsqr():
    rng = np.random.default_rng(7)
    nz, nx = 6, 9
    inv = np.arange(1, nz * nx + 1).reshape(nz, nx)
    nparm = nz * nx
    ndata = 80
    G = sp.random(ndata, nparm, density=0.2, random_state=3, format="csr")
    r = rng.normal(size=ndata)
    lap = Laplacian(inv, nparm)
    w, wd = 0.3, 0.1
    gtg = np.ones(nparm)  # nothing masked
    dm, it, *_ = run_lsqr(G, r, gtg, 1e-6, nparm, wd, wd, w, w, np.zeros(nparm), lap, max_iter=5000, atol=1e-13)
    L = sp.csr_matrix(np.array([lap(e) for e in np.eye(nparm)]).T)
    A = sp.vstack([G, w * L, wd * sp.eye(nparm)]).tocsr()
    b = np.concatenate([r, np.zeros(2 * nparm)])
    ref = spla.lsqr(A, b, atol=1e-14, btol=1e-14, iter_lim=20000)[0]
    assert np.max(np.abs(dm - ref)) < 1e-8 * max(1.0, np.max(np.abs(ref)))


def test_laplacian_stencil_closed_form():
    inv = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 0]])
    lap = Laplacian(inv, 7)
    v = np.arange(1.0, 8.0)
    # node 4 (centre): neighbours 3, 5, 1, 7 -> 3+5+1+7 - 4*4
    out = lap(v)
    assert out[3] == (3 + 5 + 1 + 7) - 4 * 4
    # node 1 (top row, x=1): neighbours x+1=2, z+1=4 -> 2+4 - 2*1
    assert out[0] == (2 + 4) - 2 * 1
    # node 6 (bottom left): neighbours x+1=7, z-1=3 -> 7+3 - 2*6
    assert out[5] == (7 + 3) - 2 * 6


GAUSS = os.path.join(ROOT, "tests", "golden", "ref_lsqr_gauss.npz")


def gauss_inputs(Z):
    from oracle.lsqr_oracle import gauss_matrices, iface_gauss_matrix
    ifaces = [dict(start=int(s), x=Z[f"iface_x{k}"]) for k, s in enumerate(Z["iface_start"])]
    return ifaces, gauss_matrices, iface_gauss_matrix


@pytest.mark.parametrize("stype", [2, 3])
def test_gaussian_operators_match_reference(stype):
    """The oracle's R_h, R_v and R_if against the matrices the reference built
    (row layout, every R_if entry, the row sums of R_h and R_v)."""
    Z = np.load(GAUSS)
    p = f"t{stype}_"
    par = Z[p + "par"]
    corr_x, corr_z, corr_xi = par[8], par[9], par[10]
    ifaces, gauss_matrices, iface_gauss_matrix = gauss_inputs(Z)
    mv = Z["mvscale"]
    rif = iface_gauss_matrix(ifaces, corr_xi, mv)
    assert np.array_equal(rif[0], Z[p + "rif_rowptr"])
    assert np.array_equal(rif[1], Z[p + "rif_col"])
    assert np.array_equal(rif[2], Z[p + "rif_val"])
    if stype == 2:
        rh, rv = gauss_matrices(Z["vx"], Z["vz"], Z["inv_idx"], corr_x, corr_z, mv)
        for name, m in (("rh", rh), ("rv", rv)):
            assert np.array_equal(m[0], Z[p + name + "_rowptr"])
            sums = np.array([np.cumsum(m[2][a - 1:b - 1])[-1] if b > a else 0.0
                             for a, b in zip(m[0][:-1], m[0][1:])])
            assert np.array_equal(sums, Z[p + name + "_rowsum"])


@pytest.mark.parametrize("stype", [2, 3])
def test_oracle_reproduces_reference_gaussian_lsqr(stype):
    """run_lsqr with Gaussian (2) and hybrid (3) smoothing: the reference's
    model update bit for bit from its own system."""
    Z = np.load(GAUSS)
    p = f"t{stype}_"
    ndata, nparm, nnz, n_vel, st, maxit, iters = (int(x) for x in Z[p + "scal"])
    assert st == stype
    wdv, wdv_i, ws, ws_i, eps, rnorm, bnorm, test2, corr_x, corr_z, corr_xi = Z[p + "par"]
    ifaces, gauss_matrices, iface_gauss_matrix = gauss_inputs(Z)
    mv = Z["mvscale"]
    rh = rv = None
    if stype == 2:
        rh, rv = gauss_matrices(Z["vx"], Z["vz"], Z["inv_idx"], corr_x, corr_z, mv)
    rif = iface_gauss_matrix(ifaces, corr_xi, mv)
    G = sp.csr_matrix((Z["gval"], Z["gindx"] - 1, Z["rowptr"] - 1), shape=(ndata, nparm))
    lap = Laplacian(Z["inv_idx"], nparm, [(int(s), len(Z[f"iface_x{k}"])) for k, s in enumerate(Z["iface_start"])])
    dm, it, rn, bn, t2 = run_lsqr(G, Z["residual"], Z["gtg"], eps, n_vel, wdv, wdv_i, ws, ws_i, Z["dm_sum"], lap,
                                  maxit, smooth_type=stype, rh=rh, rv=rv, rif=rif)
    assert it == iters
    assert rn == rnorm and bn == bnorm and t2 == test2
    assert np.array_equal(dm, Z[p + "dm"])


def test_gaussian_recurrences_converge_to_textbook_lsqr():
    """smooth_type 2 on a small grid with stopping off: the explicit matrix
    [G; w R_h; w R_v; wd I] solved by scipy's LSQR."""
    from oracle.lsqr_oracle import _csr, gauss_matrices
    rng = np.random.default_rng(11)
    nz, nx = 5, 8
    inv = np.arange(1, nz * nx + 1).reshape(nz, nx)
    nparm = nz * nx
    vx = np.arange(nx) * 0.5
    vz = np.arange(nz) * 0.5
    mv = rng.uniform(0.2, 0.4, nparm)
    rh, rv = gauss_matrices(vx, vz, inv, 0.7, 0.6, mv)
    G = sp.random(60, nparm, density=0.25, random_state=5, format="csr")
    r = rng.normal(size=60)
    w, wd = 0.3, 0.1
    dm, it, *_ = run_lsqr(G, r, np.ones(nparm), 1e-6, nparm, wd, wd, w, w, np.zeros(nparm),
                          Laplacian(inv, nparm), max_iter=5000, atol=1e-13, smooth_type=2, rh=rh, rv=rv)
    RH, RV = _csr(*rh, nparm), _csr(*rv, nparm)
    A = sp.vstack([G, w * RH, w * RV, wd * sp.eye(nparm)]).tocsr()
    b = np.concatenate([r, np.zeros(3 * nparm)])
    ref = spla.lsqr(A, b, atol=1e-14, btol=1e-14, iter_lim=20000)[0]
    assert np.max(np.abs(dm - ref)) < 1e-8 * max(1.0, np.max(np.abs(ref)))
    # closed form of one R_h row: neighbours exp(-d^2/L^2) mv_j, diagonal -sum(beta) mv_i
    row = 2 * nx + 3  # (iz=2, ix=3), parameter 20
    a, b_ = rh[0][row] - 1, rh[0][row + 1] - 1
    cols, vals = rh[1][a:b_], rh[2][a:b_]
    assert cols[-1] == row + 1
    assert list(cols[:-1]) == [2 * nx + k + 1 for k in range(nx) if k != 3]  # kmax = int(3*0.7/0.5)+1 = 5
    betas = []
    for c, v in zip(cols[:-1], vals[:-1]):
        d = vx[(c - 1) % nx] - vx[3]
        betas.append(math.exp(-(d * d) / (0.7 * 0.7)))
        assert v == betas[-1] * mv[c - 1]
    assert vals[-1] == -math.fsum(betas) * mv[row] or abs(vals[-1] + sum(betas) * mv[row]) < 1e-15
