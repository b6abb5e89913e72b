"""FP64 numpy restatement of the aggregation multigrid that shk_solve_csr applies for
Preconditioner=mg (an arbitrary symmetric CSR matrix).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's CSR
multigrid (hierarchy sizes, iteration counts); never part of the product path.

The reference's "amg" LinSolver (src/linsolver/amg.cpp:43-166) runs amgcl's smoothed
aggregation with Gauss-Seidel relaxation under CG, set up on the host for every solve.
The library's counterpart is its own design, restated here step by step:
- aggregation: two greedy pairwise passes per level (rows in index order; an
  unaggregated row pairs with its strongest unaggregated strong neighbour, strength
  -a_ij >= BETA * max_k -a_ik, ties to the lower index, else with its strongest
  unaggregated neighbour of any negative coupling, else it stays alone), the second pass
  on the Galerkin matrix of the first, so aggregates hold up to 4 rows;
- a level whose aggregation keeps more than STALL of its rows is the coarsest (the
  over-correction ALPHA compounds over levels that barely coarsen);
- coarse matrices: Galerkin products P'AP of the piecewise-constant prolongation
  (exact zeros dropped);
- levels until at most COARSE_MAX rows, then a dense solve;
- V-cycle: one damped-Jacobi step from zero (OMEGA), restriction of the residual (sum
  over the aggregate), recursion, prolongation with over-correction ALPHA, one
  damped-Jacobi post-smoothing step: a fixed symmetric positive definite M, so plain
  PCG applies.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

BETA = 0.25
COARSE_MAX = 128
STALL = 0.8
OMEGA = 0.9
ALPHA = 1.5


def pairwise(A: sp.csr_matrix):
    """Greedy pairwise aggregation: (agg, n_coarse), agg[i] = coarse row of row i."""
    n = A.shape[0]
    indptr, indices, data = A.indptr, A.indices, A.data
    agg = np.full(n, -1, dtype=np.int64)
    nc = 0
    for i in range(n):
        if agg[i] >= 0:
            continue
        lo, hi = indptr[i], indptr[i + 1]
        cols, vals = indices[lo:hi], data[lo:hi]
        off = (cols != i) & (vals < 0)
        best, bj, weak, wj = 0.0, -1, 0.0, -1
        if off.any():
            smax = (-vals[off]).max()
            for j, v in zip(cols, vals):
                if j == i or v >= 0 or agg[j] >= 0:
                    continue
                s = -v
                if s >= BETA * smax and (s > best or (s == best and j < bj)):
                    best, bj = s, j
                if s > weak or (s == weak and j < wj):
                    weak, wj = s, j
        if bj < 0:
            bj = wj
        agg[i] = nc
        if bj >= 0:
            agg[bj] = nc
        nc += 1
    return agg, nc


def galerkin(A: sp.csr_matrix, agg, nc) -> sp.csr_matrix:
    n = A.shape[0]
    P = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, nc))
    Ac = (P.T @ A @ P).tocsr()
    Ac.eliminate_zeros()
    Ac.sort_indices()
    return Ac


def hierarchy(A: sp.csr_matrix):
    """[(A_l, agg_l)] down to the coarsest (agg of the coarsest is None)."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    levels = []
    while A.shape[0] > COARSE_MAX:
        a1, n1 = pairwise(A)
        A1 = galerkin(A, a1, n1)
        a2, n2 = pairwise(A1)
        agg = a2[a1]
        if n2 > STALL * A.shape[0]:
            break
        levels.append((A, agg))
        A = galerkin(A1, a2, n2)  # = P'AP of the composed aggregation
    levels.append((A, None))
    return levels


def vcycle(levels, b, l=0):
    A, agg = levels[l]
    if agg is None:
        return np.linalg.solve(A.toarray(), b) if A.shape[0] else b
    dinv = 1.0 / A.diagonal()
    x = OMEGA * dinv * b
    t = b - A @ x
    nc = levels[l + 1][0].shape[0]
    bc = np.bincount(agg, weights=t, minlength=nc)
    xc = vcycle(levels, bc, l + 1)
    x = x + ALPHA * xc[agg]
    return x + OMEGA * dinv * (b - A @ x)


def pcg(A, b, tol=1e-6, max_iterations=2000, levels=None):
    """CG preconditioned by the V-cycle, the reference's max-norm stop (pcg_solver.h:258)."""
    levels = levels or hierarchy(A)
    A0 = levels[0][0]
    x = np.zeros_like(b)
    r = b.copy()
    bmax = np.abs(b).max()
    if bmax == 0:
        return x, 0
    z = vcycle(levels, r)
    s = z.copy()
    rho = r @ z
    for it in range(1, max_iterations + 1):
        q = A0 @ s
        a = rho / (s @ q)
        x += a * s
        r -= a * q
        if np.abs(r).max() <= tol * bmax:
            return x, it
        z = vcycle(levels, r)
        rn = r @ z
        s = z + (rn / rho) * s
        rho = rn
    return x, max_iterations
