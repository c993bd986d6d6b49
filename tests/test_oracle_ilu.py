"""Pins of the oracle's block ILU(0) (Ifpack "ILU" at fill level 0,
belos_linear_solver_factory.hpp:141-160).  CPU only.

What the mathematics fixes:
  - ILU(0) reproduces A on its own pattern: (L U)_ij = A_ij for every block
    (i, j) of the pattern (L unit block-lower, U block-upper);
  - on a block-tridiagonal matrix (a pattern closed under elimination) ILU(0)
    is the exact block LU, so its application is A^-1;
  - GMRES with it converges on the assembled channel system where block Jacobi
    stalls.
"""
import numpy as np
import pytest

import orc
from helpers import rel

from gales_mc import cases


def _dense_factors(rowptr, colidx, lu, nn, n):
    L = np.eye(nn * n)
    U = np.zeros((nn * n, nn * n))
    for i in range(nn):
        for t in range(rowptr[i], rowptr[i + 1]):
            j = colidx[t]
            blk = lu[t * n * n:(t + 1) * n * n].reshape(n, n)
            if j < i:
                L[i * n:(i + 1) * n, j * n:(j + 1) * n] = blk
            else:
                U[i * n:(i + 1) * n, j * n:(j + 1) * n] = blk
    return L, U


def test_ilu0_matches_the_matrix_on_its_pattern():
    mesh, P, I, Pst = cases.config(0)
    A = orc.Assembler(P, mesh)
    vals, _ = A.assemble(I, Pst)
    # a corner of the system keeps the dense check small: the first 150 nodes
    n = A.n
    m = 150
    rp, ci = [0], []
    v = []
    for i in range(m):
        for t in range(A.rowptr[i], A.rowptr[i + 1]):
            if A.colidx[t] < m:
                ci.append(A.colidx[t])
                v.append(vals.reshape(-1, n, n)[t])
        rp.append(len(ci))
    rp, ci, v = np.array(rp), np.array(ci), np.array(v)
    lu, dinv = orc.bilu0(rp, ci, v, n)
    L, U = _dense_factors(rp, ci, lu, m, n)
    LU = L @ U
    Ad = np.zeros_like(LU)
    for i in range(m):
        for t in range(rp[i], rp[i + 1]):
            j = ci[t]
            Ad[i * n:(i + 1) * n, j * n:(j + 1) * n] = v[t]
            blk = LU[i * n:(i + 1) * n, j * n:(j + 1) * n]
            assert rel(blk, v[t]) < 1e-12 or np.abs(blk - v[t]).max() < 1e-12 * np.abs(v).max()
    # and it is not the full LU (fill outside the pattern is dropped)
    assert rel(LU, Ad) > 1e-6
    # dinv is the inverse of U's diagonal blocks
    for i in range(m):
        Uii = U[i * n:(i + 1) * n, i * n:(i + 1) * n]
        assert rel(dinv[i * n * n:(i + 1) * n * n].reshape(n, n) @ Uii, np.eye(n)) < 1e-10


def test_ilu0_is_exact_on_a_block_tridiagonal_matrix():
    rng = np.random.default_rng(4)
    nn, n = 40, 5
    rp, ci, v = [0], [], []
    for i in range(nn):
        for j in (i - 1, i, i + 1):
            if 0 <= j < nn:
                ci.append(j)
                blk = rng.standard_normal((n, n))
                if j == i:
                    blk += 8.0 * np.eye(n)
                v.append(blk)
        rp.append(len(ci))
    rp, ci, v = np.array(rp), np.array(ci), np.array(v)
    Ad = np.zeros((nn * n, nn * n))
    for i in range(nn):
        for t in range(rp[i], rp[i + 1]):
            Ad[i * n:(i + 1) * n, ci[t] * n:(ci[t] + 1) * n] = v[t]
    lu, dinv = orc.bilu0(rp, ci, v, n)
    r = rng.standard_normal(nn * n)
    x = orc.bilu0_apply(rp, ci, lu, dinv, n, r)
    assert rel(x, np.linalg.solve(Ad, r)) < 1e-12


def test_gmres_with_ilu0_converges_where_jacobi_stalls():
    mesh, P, I, Pst = cases.config(1, scale=16)
    A = orc.Assembler(P, mesh)
    vals, rhs = A.assemble(I, Pst)
    _, it_j, rr_j = A.gmres(vals, rhs, restart=50, max_iters=300, rel_tol=1e-10, precond=2)
    x, it_i, rr_i = A.gmres(vals, rhs, restart=50, max_iters=300, rel_tol=1e-10, precond=3)
    assert rr_i < 1e-9 and it_i < 300
    assert rr_j > 1e-6
    y = A.spmv(vals, x)
    assert rel(y, rhs) < 1e-9
