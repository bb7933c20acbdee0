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
